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
#include "galerkin_kernels.hpp"
#include "mg_kernels.hpp"

namespace icarat_hip {

namespace {

inline unsigned blocks_for(long long n, int nt) { return (unsigned)((n + nt - 1) / nt); }

// XCD-aware tile order for a 1-D grid over nbx x nby x nbz tiles: workgroups go
// to the 8 XCDs round-robin by id, so each XCD is given a contiguous range of
// tiles (x fastest) and neighbouring tiles share their halos in its L2.
__device__ __forceinline__ void xcd_tile(int nbx, int nby, int &bx, int &by, int &bz) {
  const unsigned nt = gridDim.x, b = blockIdx.x;
  const unsigned q = nt / 8, rem = nt % 8, xcd = b % 8, slot = b / 8;
  const unsigned T = (xcd < rem ? xcd * (q + 1) : rem * (q + 1) + (xcd - rem) * q) + slot;
  bx = (int)(T % nbx);
  by = (int)((T / nbx) % nby);
  bz = (int)(T / ((unsigned)nbx * nby));
}

// Restriction and prolongation are tiled through LDS.  A restriction block
// owns RTX x RTY x RTZ coarse nodes and stages the fine region under them (one
// fine node of halo on each side) with the residual formed on load, so each
// fine value is read from HBM about once rather than by every coarse node that
// weighs it.  A prolongation block owns PTX x PTY x PTZ fine nodes and stages
// the coarse nodes around them.  Lattice nodes outside the grid enter as 0,
// which is the masked transfer's value there.
constexpr int RTX = 16, RTY = 4, RTZ = 4;
constexpr int RFX = 2 * RTX + 1, RFY = 2 * RTY + 1, RFZ = 2 * RTZ + 1, RFN = RFX * RFY * RFZ;
constexpr int PTX = 32, PTY = 4, PTZ = 2;
constexpr int PCX = PTX / 2 + 1, PCY = PTY / 2 + 1, PCZ = PTZ / 2 + 1, PCN = PCX * PCY * PCZ;

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
    if constexpr (MODE == MG_JACOBI) v = xc[c] + omega * dinv[q] * (b[q] - y[c]);
    else v = b[q] - y[c];
    out[q] = ((m >> c) & 1u) ? v : 0.0;
  }
}

// bc = R (rf - yf) (or R rf), masked; with dinv_c also xc = omega dinv_c bc
template <bool HAS_Y>
__global__ __launch_bounds__(256) void k_mg_restrict(Lattice Lf, Lattice Lc,
                                                     const double *__restrict__ rf,
                                                     const double *__restrict__ yf,
                                                     const uint8_t *__restrict__ fmask_c,
                                                     double *__restrict__ bc,
                                                     const double *__restrict__ dinv_c,
                                                     double omega, double *__restrict__ xc,
                                                     long long sfs, long long scs) {
  __shared__ double sv[3][RFN];
  // load case blockIdx.y: its vectors start at blockIdx.y x the level's stride
  rf += blockIdx.y * sfs;
  if (HAS_Y) yf += blockIdx.y * sfs;
  bc += blockIdx.y * scs;
  if (xc) xc += blockIdx.y * scs;
  int bx, by, bz;
  xcd_tile((Lc.NX + RTX - 1) / RTX, (Lc.NY + RTY - 1) / RTY, bx, by, bz);
  const int I0 = bx * RTX, J0 = by * RTY, K0 = bz * RTZ;
  const int fx0 = 2 * I0 - 1, fy0 = 2 * J0 - 1, fz0 = 2 * K0 - 1;
  // all loads of the staging are issued before any is stored (load phase into
  // registers, then the LDS phase)
  constexpr int NK = (RFN + 255) / 256;
  double v[NK][3];
#pragma unroll
  for (int k = 0; k < NK; ++k) {
    const int t = threadIdx.x + 256 * k;
    const int hx = t % RFX, hy = (t / RFX) % RFY, hz = t / (RFX * RFY);
    const int gx = fx0 + hx, gy = fy0 + hy, gz = fz0 + hz;
    const bool in = t < RFN && gx >= 0 && gx < Lf.NX && gy >= 0 && gy < Lf.NY && gz >= 0 && gz < Lf.NZ;
    const long long f = in ? gx + (long long)Lf.NX * (gy + (long long)Lf.NY * gz) : 0;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const double a = in ? rf[c * Lf.Np + f] : 0.0;
      const double b = (HAS_Y && in) ? yf[c * Lf.Np + f] : 0.0;
      v[k][c] = a - b;
    }
  }
#pragma unroll
  for (int k = 0; k < NK; ++k) {
    const int t = threadIdx.x + 256 * k;
    if (t < RFN) {
#pragma unroll
      for (int c = 0; c < 3; ++c) sv[c][t] = v[k][c];
    }
  }
  __syncthreads();
  const int tx = threadIdx.x % RTX, ty = (threadIdx.x / RTX) % RTY, tz = threadIdx.x / (RTX * RTY);
  const int I = I0 + tx, J = J0 + ty, K = K0 + tz;
  if (I >= Lc.NX || J >= Lc.NY || K >= Lc.NZ) return;
  double s[3] = {0.0, 0.0, 0.0};
#pragma unroll
  for (int dz = -1; dz <= 1; ++dz)
#pragma unroll
    for (int dy = -1; dy <= 1; ++dy)
#pragma unroll
      for (int dx = -1; dx <= 1; ++dx) {
        const double w = (dx == 0 ? 1.0 : 0.5) * (dy == 0 ? 1.0 : 0.5) * (dz == 0 ? 1.0 : 0.5);
        const int h = (2 * tx + 1 + dx) + RFX * ((2 * ty + 1 + dy) + RFY * (2 * tz + 1 + dz));
#pragma unroll
        for (int c = 0; c < 3; ++c) s[c] += w * sv[c][h];
      }
  const long long n = I + (long long)Lc.NX * (J + (long long)Lc.NY * K);
  const unsigned m = fmask_c[n];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const double v = ((m >> c) & 1u) ? s[c] : 0.0;
    bc[c * Lc.Np + n] = v;
    if (dinv_c) xc[c * Lc.Np + n] = omega * dinv_c[c * Lc.Np + n] * v;
  }
}

// xf += P xc (ADD) or xf = P xc (trilinear), masked by the fine free DOFs
// (xc2 != nullptr: the coarse values are xc + xc2, the W-cycle's two visits)
template <bool ADD, class TF = double>
__global__ __launch_bounds__(256) void k_mg_prolong_add(Lattice Lf, Lattice Lc,
                                                        const double *__restrict__ xc,
                                                        const double *__restrict__ xc2,
                                                        const uint8_t *__restrict__ fmask_f,
                                                        TF *__restrict__ xf, long long sfs,
                                                        long long scs) {
  __shared__ double sc[3][PCN];
  xc += blockIdx.y * scs;
  if (xc2) xc2 += blockIdx.y * scs;
  xf += blockIdx.y * sfs;
  int bx, by, bz;
  xcd_tile((Lf.NX + PTX - 1) / PTX, (Lf.NY + PTY - 1) / PTY, bx, by, bz);
  const int fx0 = bx * PTX, fy0 = by * PTY, fz0 = bz * PTZ;
  const int I0 = fx0 / 2, J0 = fy0 / 2, K0 = fz0 / 2;
  const int tx = threadIdx.x % PTX, ty = (threadIdx.x / PTX) % PTY, tz = threadIdx.x / (PTX * PTY);
  const int x = fx0 + tx, y = fy0 + ty, z = fz0 + tz;
  const bool own = x < Lf.NX && y < Lf.NY && z < Lf.NZ;
  const long long n = own ? x + (long long)Lf.NX * (y + (long long)Lf.NY * z) : 0;
  // the fine values and mask are loaded before the coarse staging, so both
  // round trips overlap
  double xo[3] = {0.0, 0.0, 0.0};
  unsigned m = 0;
  if (own) {
    m = fmask_f[n];
    if constexpr (ADD) {
#pragma unroll
      for (int c = 0; c < 3; ++c) xo[c] = (double)xf[c * Lf.Np + n];
    }
  }
  for (int t = threadIdx.x; t < PCN; t += 256) {
    const int hx = t % PCX, hy = (t / PCX) % PCY, hz = t / (PCX * PCY);
    const int gx = I0 + hx, gy = J0 + hy, gz = K0 + hz;
    double v[3] = {0.0, 0.0, 0.0};
    if (gx < Lc.NX && gy < Lc.NY && gz < Lc.NZ) {
      const long long q = gx + (long long)Lc.NX * (gy + (long long)Lc.NY * gz);
#pragma unroll
      for (int c = 0; c < 3; ++c) v[c] = xc2 ? xc[c * Lc.Np + q] + xc2[c * Lc.Np + q] : xc[c * Lc.Np + q];
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) sc[c][t] = v[c];
  }
  __syncthreads();
  if (!own) return;
  // local coarse index of the lower neighbour and the 1-D weights (an even
  // fine coordinate sits on a coarse node: weights 1, 0; odd: 1/2, 1/2)
  const int lx = tx >> 1, ly = ty >> 1, lz = tz >> 1;
  const double wx1 = (tx & 1) ? 0.5 : 0.0, wy1 = (ty & 1) ? 0.5 : 0.0, wz1 = (tz & 1) ? 0.5 : 0.0;
  const double wx[2] = {1.0 - wx1, wx1}, wy[2] = {1.0 - wy1, wy1}, wz[2] = {1.0 - wz1, wz1};
  double s[3] = {0.0, 0.0, 0.0};
#pragma unroll
  for (int c2 = 0; c2 < 2; ++c2)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int a = 0; a < 2; ++a) {
        const double w = wx[a] * wy[b] * wz[c2];
        const int h = (lx + a) + PCX * ((ly + b) + PCY * (lz + c2));
#pragma unroll
        for (int c = 0; c < 3; ++c) s[c] += w * sc[c][h];
      }
#pragma unroll
  for (int c = 0; c < 3; ++c)
    if ((m >> c) & 1u) xf[c * Lf.Np + n] = (TF)(xo[c] + s[c]);
}

// x = omega dinv b in FP32 (the first sweep from zero of an FP32 level)
__global__ void k_mg_jacobi0_f32(long long n, double omega, const double *__restrict__ dinv,
                                 const double *__restrict__ b, float *__restrict__ x) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i < n) x[i] = (float)(omega * dinv[i] * b[i]);
}

// coarse element from its 8 children: arithmetic mean (MODE 0), geometric mean
// (1), quadratic mean (3) or maximum (4) of the values
template <int MODE>
__global__ void k_mg_young(Lattice Lf, Lattice Lc, const double *__restrict__ yf,
                           double *__restrict__ yc) {
  const long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (e >= Lc.nelx) return;
  const int I = (int)(e % Lc.nx), J = (int)((e / Lc.nx) % Lc.ny), K = (int)(e / ((long long)Lc.nx * Lc.ny));
  double s = 0.0;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const int x = 2 * I + (k & 1), y = 2 * J + ((k >> 1) & 1), z = 2 * K + (k >> 2);
    const double v = yf[x + (long long)Lf.nx * (y + (long long)Lf.ny * z)];
    if constexpr (MODE == 1) s += log(v);
    else if constexpr (MODE == 3) s += v * v;
    else if constexpr (MODE == 4) s = fmax(s, v);
    else s += v;
  }
  if constexpr (MODE == 1) yc[e] = exp(0.125 * s);
  else if constexpr (MODE == 3) yc[e] = sqrt(0.125 * s);
  else if constexpr (MODE == 4) yc[e] = s;
  else yc[e] = 0.125 * s;
}

// x = omega dinv b (first sweep from zero) or x += omega dinv (b - y)
__global__ void k_mg_jacobi(long long n, double omega, const double *__restrict__ dinv,
                            const double *__restrict__ b, const double *__restrict__ y,
                            double *__restrict__ x) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  // load case blockIdx.y (stride n; dinv is the level's, shared)
  const long long o = blockIdx.y * n;
  b += o;
  x += o;
  if (y) y += o;
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
                                                       const double *xin, double *xout,
                                                       const double *__restrict__ q,
                                                       double *part) {
  __shared__ double sh[8];
  double zr = 0.0, zq = 0.0;
  const long long stride = (long long)gridDim.x * 256;
  for (long long i = blockIdx.x * 256LL + threadIdx.x; i < n; i += stride) {
    const double bi = b[i];
    const double xi = xin[i] + omega * dinv[i] * (bi - y[i]);
    xout[i] = xi;
    zr += xi * bi;
    zq += xi * q[i];
  }
  double v[2] = {zr, zq};
  store_partials<256, 2>(v, part, sh);
}

template <int SYM, int MODE>
__global__ __launch_bounds__(256) void k_mg_direct(Lattice L, const double *__restrict__ c_KE0,
                                                   KeVals kv, const double *__restrict__ young,
                                                   const uint8_t *__restrict__ fmask,
                                                   const double *__restrict__ x,
                                                   const double *__restrict__ b,
                                                   const double *__restrict__ dinv, double omega,
                                                   double *__restrict__ out, long long cs) {
  const long long n = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (n >= L.numnp) return;
  x += blockIdx.y * cs;
  b += blockIdx.y * cs;
  out += blockIdx.y * cs;
  double y[3], xc[3];
  direct_node<SYM>(L, c_KE0, kv, young, x, n, y, xc);
  direct_out<MODE>(L.Np, n, fmask[n], y, xc, b, dinv, omega, out);
}

// The start of the W-cycle's second visit of a direct level, in one pass over
// the first visit's result xc: u = xc (kept for the coarse correction),
// r = b - K xc (masked) and, with x0, the second visit's first sweep from zero
// x0 = omega dinv r.  xc may not alias u, r or x0.
template <int SYM>
__global__ __launch_bounds__(256) void k_mg_direct_wres(Lattice L, const double *__restrict__ c_KE0,
                                                        KeVals kv, const double *__restrict__ young,
                                                        const uint8_t *__restrict__ fmask,
                                                        const double *__restrict__ xc,
                                                        const double *__restrict__ b,
                                                        const double *__restrict__ dinv,
                                                        double omega, double *__restrict__ u,
                                                        double *__restrict__ r,
                                                        double *__restrict__ x0, long long cs) {
  const long long n = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (n >= L.numnp) return;
  xc += blockIdx.y * cs;
  b += blockIdx.y * cs;
  u += blockIdx.y * cs;
  r += blockIdx.y * cs;
  if (x0) x0 += blockIdx.y * cs;
  double y[3], xn[3];
  direct_node<SYM>(L, c_KE0, kv, young, xc, n, y, xn);
  const unsigned m = fmask[n];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const long long q = c * L.Np + n;
    u[q] = xn[c];
    const double v = ((m >> c) & 1u) ? b[q] - y[c] : 0.0;
    r[q] = v;
    if (x0) x0[q] = omega * dinv[q] * v;
  }
}

// The coarse correction and the post-smoothing sweep of a direct level in one
// pass: xp = x + P xc (masked by the fine free DOFs, as k_mg_prolong_add<true>
// forms it), out = xp + omega dinv (b - K xp) (masked; xp itself is not
// stored).  A block owns QTX x QTY x QTZ fine nodes; it stages the coarse nodes
// under them, then xp on the tile and one node of halo, then contracts from LDS.
constexpr int QTX = 16, QTY = 4, QTZ = 4;
constexpr int QHX = QTX + 2, QHY = QTY + 2, QHZ = QTZ + 2, QHN = QHX * QHY * QHZ;
// Halo fine node g reads coarse nodes g / 2 and g / 2 + 1 (the second with
// weight 0 when g is even): local coarse indices 0 .. QT / 2 + 2.
constexpr int QCX = QTX / 2 + 3, QCY = QTY / 2 + 3, QCZ = QTZ / 2 + 3, QCN = QCX * QCY * QCZ;

template <int SYM>
__global__ __launch_bounds__(256) void k_mg_prolong_jacobi(
    Lattice Lf, Lattice Lc, const double *__restrict__ c_KE0, KeVals kv,
    const double *__restrict__ young, const uint8_t *__restrict__ fmask,
    const double *__restrict__ xcoarse, const double *__restrict__ xcoarse2,
    const double *__restrict__ x, const double *__restrict__ b, const double *__restrict__ dinv,
    double omega, double *__restrict__ out, long long sfs, long long scs) {
  __shared__ double sc[3][QCN];
  __shared__ double su[3][QHN];
  // load case blockIdx.z / (z tiles): its vectors at case x the level's stride
  const int nbz = (Lf.NZ + QTZ - 1) / QTZ, lc = blockIdx.z / nbz;
  xcoarse += lc * scs;
  if (xcoarse2) xcoarse2 += lc * scs;
  x += lc * sfs;
  b += lc * sfs;
  out += lc * sfs;
  const int fx0 = blockIdx.x * QTX, fy0 = blockIdx.y * QTY, fz0 = (blockIdx.z % nbz) * QTZ;
  // coarse node C0 + h covers fine halo nodes from 2 (C0 + h) - 1 to 2 (C0 + h) + 1
  const int cx0 = fx0 / 2 - 1, cy0 = fy0 / 2 - 1, cz0 = fz0 / 2 - 1;
  const int tx = threadIdx.x % QTX, ty = (threadIdx.x / QTX) % QTY, tz = threadIdx.x / (QTX * QTY);
  const int i = fx0 + tx, j = fy0 + ty, k = fz0 + tz;
  const bool own = i < Lf.NX && j < Lf.NY && k < Lf.NZ;
  const long long n = own ? i + (long long)Lf.NX * (j + (long long)Lf.NY * k) : 0;
  // every global load is issued before the first barrier: the coarse nodes,
  // the halo's x and mask, and the owned node's elements, b and dinv
  constexpr int NC = (QCN + 255) / 256, NH = (QHN + 255) / 256;
  double cv[NC][3];
#pragma unroll
  for (int r = 0; r < NC; ++r) {
    const int t = threadIdx.x + 256 * r;
    const int gx = cx0 + t % QCX, gy = cy0 + (t / QCX) % QCY, gz = cz0 + t / (QCX * QCY);
    const bool in = t < QCN && gx >= 0 && gx < Lc.NX && gy >= 0 && gy < Lc.NY && gz >= 0 && gz < Lc.NZ;
    const long long q = in ? gx + (long long)Lc.NX * (gy + (long long)Lc.NY * gz) : 0;
#pragma unroll
    for (int c = 0; c < 3; ++c)
      cv[r][c] = in ? (xcoarse2 ? xcoarse[c * Lc.Np + q] + xcoarse2[c * Lc.Np + q]
                                : xcoarse[c * Lc.Np + q])
                    : 0.0;
  }
  double hv[NH][3];
  unsigned hm[NH];
  bool hin[NH];
#pragma unroll
  for (int r = 0; r < NH; ++r) {
    const int t = threadIdx.x + 256 * r;
    const int gx = fx0 - 1 + t % QHX, gy = fy0 - 1 + (t / QHX) % QHY, gz = fz0 - 1 + t / (QHX * QHY);
    hin[r] = t < QHN && gx >= 0 && gx < Lf.NX && gy >= 0 && gy < Lf.NY && gz >= 0 && gz < Lf.NZ;
    const long long m = hin[r] ? gx + (long long)Lf.NX * (gy + (long long)Lf.NY * gz) : 0;
    hm[r] = hin[r] ? fmask[m] : 0u;
#pragma unroll
    for (int c = 0; c < 3; ++c) hv[r][c] = hin[r] ? x[c * Lf.Np + m] : 0.0;
  }
  double E[8], bo[3], dv[3];
  unsigned mo = 0;
#pragma unroll
  for (int s = 0; s < 8; ++s) {
    const int ex = i + (s & 1) - 1, ey = j + ((s >> 1) & 1) - 1, ez = k + ((s >> 2) & 1) - 1;
    const bool in = own && ex >= 0 && ex < Lf.nx && ey >= 0 && ey < Lf.ny && ez >= 0 && ez < Lf.nz;
    E[s] = in ? young[ex + (long long)Lf.nx * (ey + (long long)Lf.ny * ez)] : 0.0;
  }
  if (own) {
    mo = fmask[n];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      bo[c] = b[c * Lf.Np + n];
      dv[c] = dinv[c * Lf.Np + n];
    }
  }
#pragma unroll
  for (int r = 0; r < NC; ++r) {
    const int t = threadIdx.x + 256 * r;
    if (t < QCN) {
#pragma unroll
      for (int c = 0; c < 3; ++c) sc[c][t] = cv[r][c];
    }
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < NH; ++r) {
    const int t = threadIdx.x + 256 * r;
    if (t >= QHN) continue;
    double v[3] = {0.0, 0.0, 0.0};
    if (hin[r]) {
      const int gx = fx0 - 1 + t % QHX, gy = fy0 - 1 + (t / QHX) % QHY, gz = fz0 - 1 + t / (QHX * QHY);
      // lower coarse neighbour (local) and 1-D weights, as in k_mg_prolong_add
      const int lx = (gx >> 1) - cx0, ly = (gy >> 1) - cy0, lz = (gz >> 1) - cz0;
      const double wx1 = (gx & 1) ? 0.5 : 0.0, wy1 = (gy & 1) ? 0.5 : 0.0, wz1 = (gz & 1) ? 0.5 : 0.0;
      const double wx[2] = {1.0 - wx1, wx1}, wy[2] = {1.0 - wy1, wy1}, wz[2] = {1.0 - wz1, wz1};
      double s[3] = {0.0, 0.0, 0.0};
#pragma unroll
      for (int c2 = 0; c2 < 2; ++c2)
#pragma unroll
        for (int bb = 0; bb < 2; ++bb)
#pragma unroll
          for (int a = 0; a < 2; ++a) {
            const double w = wx[a] * wy[bb] * wz[c2];
            const int h = (lx + a) + QCX * ((ly + bb) + QCY * (lz + c2));
#pragma unroll
            for (int c = 0; c < 3; ++c) s[c] += w * sc[c][h];
          }
#pragma unroll
      for (int c = 0; c < 3; ++c) v[c] = ((hm[r] >> c) & 1u) ? hv[r][c] + s[c] : hv[r][c];
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) su[c][t] = v[c];
  }
  __syncthreads();
  if (!own) return;
  double U[27][3];
#pragma unroll
  for (int dz = -1; dz <= 1; ++dz)
#pragma unroll
    for (int dy = -1; dy <= 1; ++dy)
#pragma unroll
      for (int dx = -1; dx <= 1; ++dx) {
        const int h = (tx + 1 + dx) + QHX * ((ty + 1 + dy) + QHY * (tz + 1 + dz));
#pragma unroll
        for (int c = 0; c < 3; ++c) U[nb(dx, dy, dz)][c] = su[c][h];
      }
  double y[3];
  node_contract<SYM>(c_KE0, kv, U, E, y);
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const double v = U[13][c] + omega * dv[c] * (bo[c] - y[c]);
    out[c * Lf.Np + n] = ((mo >> c) & 1u) ? v : 0.0;
  }
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

// diagonal of the free DOFs times (1 + shift)
__global__ void k_mg_dense_shift(Lattice L, const uint8_t *__restrict__ fmask, double shift,
                                 double *__restrict__ A) {
  const long long n = 3 * L.numnp;
  const long long r = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (r >= n) return;
  if ((fmask[r % L.numnp] >> (r / L.numnp)) & 1u) A[r * n + r] *= 1.0 + shift;
}

// x = A^-1 b on the coarsest lattice with the symmetric A^-1 (3 numnp square,
// row and column j = c numnp + n for node n, component c) read once on and
// below its diagonal, half the bytes of a full product.  The factorization
// leaves A^-1 column-major in its lower triangle; k_mg_dense_pack copies that
// triangle into DT x DT tiles (I, J), I >= J, stored one after another (tile
// t = I (I + 1) / 2 + J) and column-major inside, so that one tile is one
// contiguous 32 KB read.  Tile (I, J) gives the rows I the products A_IJ b_J
// and, for I > J, the rows J the products A_IJ^T b_I.  Each tile stores its
// two partial vectors in the slots (I, J) and (J, I) of a dense_tiles^2 x DT
// partial array; k_mg_dense_sum adds the dense_tiles slots of every row in a
// fixed order (deterministic, no atomics).
constexpr int DT = 64;
__host__ __device__ __forceinline__ long long dense_tiles(long long n) { return (n + DT - 1) / DT; }
__device__ __forceinline__ void dense_tile_ij(int t, int &I, int &J) {
  I = (int)((sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
  while ((I + 1) * (I + 2) / 2 <= t) ++I;
  while (I * (I + 1) / 2 > t) --I;
  J = t - I * (I + 1) / 2;
}

// packed tile t, element (r, c) at t DT^2 + c DT + r = A(I DT + r, J DT + c),
// taken from the lower triangle (the upper half of a diagonal tile from its
// mirror), zero outside the n x n matrix
__global__ void k_mg_dense_pack(int n, const double *__restrict__ A, double *__restrict__ P,
                                long long total) {
  const long long q = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (q >= total) return;
  int I, J;
  dense_tile_ij((int)(q / (DT * DT)), I, J);
  const int r = (int)(q % DT), c = (int)((q / DT) % DT);
  const int i = I * DT + r, j = J * DT + c;
  double v = 0.0;
  if (i < n && j < n) v = i >= j ? A[i + (long long)j * n] : A[j + (long long)i * n];
  P[q] = v;
}

__device__ __forceinline__ double dense_rhs(const Lattice &L, const double *__restrict__ b, int j,
                                            int n) {
  const int np1 = (int)L.numnp;
  return j < n ? b[(long long)(j / np1) * L.Np + j % np1] : 0.0;
}

// One packed tile per block of 512 threads, staged through LDS: each wave
// reads whole 512-byte tile columns (the tile is one contiguous 32 KB read).
// The first 256 threads then form the row sums A_IJ b_J and, for I > J, the
// other 256 the column sums A_IJ^T b_I, four threads per row (per column),
// each adding 16 consecutive products, the four partials added in a fixed
// tree.  One tile per block (rather than a grid-stride loop over tiles) lets
// the dispatcher start a new tile's reads as soon as any block on a CU
// finishes; the tile's LDS limits a CU to 4 blocks, and 512 threads keep twice
// the reads in flight that 256 would.
__global__ __launch_bounds__(512) void k_mg_dense_symv(Lattice L, const double *__restrict__ P,
                                                       const double *__restrict__ b,
                                                       double *__restrict__ part, int nc,
                                                       long long bs, long long ps) {
  __shared__ double tile[DT][DT + 1];
  __shared__ double sbi[DT], sbj[DT];
  const int n = (int)(3 * L.numnp), nt = (int)dense_tiles(n), ntiles = nt * (nt + 1) / 2;
  const int t = blockIdx.x;
  if (t >= ntiles) return;
  // tile t -> (I, J) by a scalar search (t is uniform across the block)
  int I = 0;
  while ((I + 1) * (I + 2) / 2 <= t) ++I;
  const int J = t - I * (I + 1) / 2;
  {
    // packed element (r, c) at c DT + r: wave w reads columns w + 8 k
    const double *__restrict__ tp = P + (long long)t * DT * DT;
    const int r = threadIdx.x & 63, w = threadIdx.x >> 6;
    double a[DT / 8];
#pragma unroll
    for (int k = 0; k < DT / 8; ++k) a[k] = tp[(w + 8 * k) * DT + r];
#pragma unroll
    for (int k = 0; k < DT / 8; ++k) tile[r][w + 8 * k] = a[k];
  }
  const bool cols = threadIdx.x >= 256;  // column sums (uniform per wave)
  const int e = (threadIdx.x & 255) >> 2, q = threadIdx.x & 3;
  // then every right-hand side (load case lc at lc x bs, its partials at lc x ps)
  for (int lc = 0; lc < nc; ++lc) {
    if (threadIdx.x < DT) {
      sbi[threadIdx.x] = dense_rhs(L, b + lc * bs, I * DT + threadIdx.x, n);
      sbj[threadIdx.x] = dense_rhs(L, b + lc * bs, J * DT + threadIdx.x, n);
    }
    __syncthreads();
    double u = 0.0;
    if (!cols) {
#pragma unroll
      for (int k = 0; k < DT / 4; ++k) {
        const int o = q * (DT / 4) + k;
        u += tile[e][o] * sbj[o];
      }
    } else if (I > J) {
#pragma unroll
      for (int k = 0; k < DT / 4; ++k) {
        const int o = q * (DT / 4) + k;
        u += tile[o][e] * sbi[o];
      }
    }
    u += __shfl_xor(u, 1);
    u += __shfl_xor(u, 2);
    if (q == 0) {
      double *pc = part + lc * ps;
      if (!cols) pc[((long long)I * nt + J) * DT + e] = u;
      else if (I > J) pc[((long long)J * nt + I) * DT + e] = u;
    }
    __syncthreads();  // the b slices are refilled for the next case
  }
}

// x rows of block I: the dense_tiles partial slots summed (a quarter of the
// slots per thread, then the four quarters in order); load case blockIdx.y
__global__ __launch_bounds__(256) void k_mg_dense_sum(Lattice L, const double *__restrict__ part,
                                                      double *__restrict__ x, long long ps,
                                                      long long xs) {
  __shared__ double sh[4][DT];
  const int n = (int)(3 * L.numnp), nt = (n + DT - 1) / DT, np1 = (int)L.numnp;
  part += blockIdx.y * ps;
  x += blockIdx.y * xs;
  const int I = blockIdx.x, r = threadIdx.x & 63, q = threadIdx.x >> 6;
  double s = 0.0;
  for (int J = q; J < nt; J += 4) s += part[((long long)I * nt + J) * DT + r];
  sh[q][r] = s;
  __syncthreads();
  if (q == 0) {
    const int i = I * DT + r;
    if (i < n) x[(long long)(i / np1) * L.Np + i % np1] = (sh[0][r] + sh[1][r]) + (sh[2][r] + sh[3][r]);
  }
}

template <int SYM>
void direct_dispatch(int mode, const Lattice &L, const KeCoef &kc, const double *young,
                     const uint8_t *fmask, const double *x, const double *b, const double *dinv,
                     double omega, double *out, hipStream_t s, int nc, long long cs) {
  const dim3 g(blocks_for(L.numnp, 256), (unsigned)nc), t(256);
  if (mode == MG_JACOBI)
    hipLaunchKernelGGL((k_mg_direct<SYM, MG_JACOBI>), g, t, 0, s, L, kc.ke, kc.vals, young, fmask,
                       x, b, dinv, omega, out, cs);
  else
    hipLaunchKernelGGL((k_mg_direct<SYM, MG_RESIDUAL>), g, t, 0, s, L, kc.ke, kc.vals, young,
                       fmask, x, b, dinv, omega, out, cs);
}

}  // namespace

void launch_mg_prolong_add_f32(const Lattice &Lf, const Lattice &Lc, const double *xc,
                               const uint8_t *fmask_f, float *xf, hipStream_t s) {
  if (Lf.NX != 2 * Lc.NX - 1 || Lf.NY != 2 * Lc.NY - 1 || Lf.NZ != 2 * Lc.NZ - 1)
    throw Error(kInvalidArgument, "prolongation: lattices are not 2x apart");
  const dim3 g((unsigned)(((Lf.NX + PTX - 1) / PTX) * ((Lf.NY + PTY - 1) / PTY) *
                          ((Lf.NZ + PTZ - 1) / PTZ)));
  hipLaunchKernelGGL((k_mg_prolong_add<true, float>), g, dim3(256), 0, s, Lf, Lc, xc, nullptr, fmask_f,
                     xf, 0LL, 0LL);
  HIPCHK(hipGetLastError());
}

void launch_mg_jacobi0_f32(long long n, double omega, const double *dinv, const double *b,
                           float *x, hipStream_t s) {
  hipLaunchKernelGGL(k_mg_jacobi0_f32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n,
                     omega, dinv, b, x);
  HIPCHK(hipGetLastError());
}

void launch_mg_restrict(const Lattice &Lf, const Lattice &Lc, const double *rf, const double *yf,
                        const uint8_t *fmask_c, double *bc, const double *dinv_c, double omega,
                        double *xc, hipStream_t s, int nc, long long sf, long long sc) {
  if (Lf.NX != 2 * Lc.NX - 1 || Lf.NY != 2 * Lc.NY - 1 || Lf.NZ != 2 * Lc.NZ - 1)
    throw Error(kInvalidArgument, "restriction: lattices are not 2x apart");
  if (nc < 1) throw Error(kInternal, "restriction: no load case");
  const dim3 g((unsigned)(((Lc.NX + RTX - 1) / RTX) * ((Lc.NY + RTY - 1) / RTY) *
                          ((Lc.NZ + RTZ - 1) / RTZ)),
               (unsigned)nc);
  if (yf)
    hipLaunchKernelGGL(k_mg_restrict<true>, g, dim3(256), 0, s, Lf, Lc, rf, yf, fmask_c, bc, dinv_c,
                       omega, xc, sf, sc);
  else
    hipLaunchKernelGGL(k_mg_restrict<false>, g, dim3(256), 0, s, Lf, Lc, rf, yf, fmask_c, bc, dinv_c,
                       omega, xc, sf, sc);
  HIPCHK(hipGetLastError());
}

void launch_mg_prolong_add(const Lattice &Lf, const Lattice &Lc, const double *xc,
                           const uint8_t *fmask_f, double *xf, hipStream_t s, bool add,
                           const double *xc2, int nc, long long sf, long long sc) {
  if (Lf.NX != 2 * Lc.NX - 1 || Lf.NY != 2 * Lc.NY - 1 || Lf.NZ != 2 * Lc.NZ - 1)
    throw Error(kInvalidArgument, "prolongation: lattices are not 2x apart");
  if (nc < 1) throw Error(kInternal, "prolongation: no load case");
  const dim3 g((unsigned)(((Lf.NX + PTX - 1) / PTX) * ((Lf.NY + PTY - 1) / PTY) *
                          ((Lf.NZ + PTZ - 1) / PTZ)),
               (unsigned)nc);
  if (add)
    hipLaunchKernelGGL(k_mg_prolong_add<true>, g, dim3(256), 0, s, Lf, Lc, xc, xc2, fmask_f, xf, sf,
                       sc);
  else
    hipLaunchKernelGGL(k_mg_prolong_add<false>, g, dim3(256), 0, s, Lf, Lc, xc, xc2, fmask_f, xf, sf,
                       sc);
  HIPCHK(hipGetLastError());
}

void launch_mg_dense_assemble(const Lattice &L, const double *ke, const double *young,
                              const uint8_t *fmask, double *A, hipStream_t s, const double *kmat) {
  const long long n = 3 * L.numnp;
  if (n > MG_DENSE_MAX) throw Error(kInvalidArgument, "lattice too large for the dense solve");
  hipLaunchKernelGGL(k_mg_dense_zero, dim3(blocks_for(n * n, 256)), dim3(256), 0, s, n * n, A);
  if (kmat) launch_gal_dense_scatter(L, kmat, A, s);
  for (int cls = 0; cls < 8 && !kmat; ++cls) {
    const int px = cls & 1, py = (cls >> 1) & 1, pz = cls >> 2;
    const long long cnt = (long long)((L.nx - px + 1) / 2) * ((L.ny - py + 1) / 2) * ((L.nz - pz + 1) / 2) * 576;
    if (cnt == 0) continue;
    hipLaunchKernelGGL(k_mg_dense_scatter, dim3(blocks_for(cnt, 256)), dim3(256), 0, s, L, ke, young,
                       px, py, pz, A);
  }
  hipLaunchKernelGGL(k_mg_dense_constrain, dim3(blocks_for(n * n, 256)), dim3(256), 0, s, L, fmask, A);
  HIPCHK(hipGetLastError());
}

void launch_mg_dense_shift(int n, const Lattice &L, const uint8_t *fmask, double shift, double *A,
                           hipStream_t s) {
  if (n != 3 * L.numnp) throw Error(kInternal, "dense shift: size mismatch");
  hipLaunchKernelGGL(k_mg_dense_shift, dim3(blocks_for(n, 256)), dim3(256), 0, s, L, fmask, shift, A);
  HIPCHK(hipGetLastError());
}

long long mg_dense_store_size(long long n) {
  const long long nt = dense_tiles(n);
  return n * n + nt * (nt + 1) / 2 * DT * DT;
}

void launch_mg_dense_pack(int n, double *A, hipStream_t s) {
  const long long nt = dense_tiles(n), total = nt * (nt + 1) / 2 * DT * DT;
  hipLaunchKernelGGL(k_mg_dense_pack, dim3(blocks_for(total, 256)), dim3(256), 0, s, n, A,
                     A + (long long)n * n, total);
  HIPCHK(hipGetLastError());
}

long long mg_dense_part_size(const Lattice &L) {
  const long long nt = dense_tiles(3 * L.numnp);
  return nt * nt * DT;
}

void launch_mg_dense_solve(const Lattice &L, const double *ainv, const double *b, double *x,
                           double *part, hipStream_t s, int nc, long long cs) {
  if (3 * L.numnp > MG_DENSE_MAX) throw Error(kInvalidArgument, "lattice too large for the dense solve");
  if (nc < 1) throw Error(kInternal, "dense solve: no load case");
  const long long n = 3 * L.numnp, nt = dense_tiles(n), ps = mg_dense_part_size(L);
  // one block per packed tile (A^-1 read once for every load case)
  const long long ntiles = nt * (nt + 1) / 2;
  hipLaunchKernelGGL(k_mg_dense_symv, dim3((unsigned)ntiles), dim3(512), 0, s, L, ainv + n * n, b,
                     part, nc, cs, ps);
  hipLaunchKernelGGL(k_mg_dense_sum, dim3((unsigned)nt, (unsigned)nc), dim3(256), 0, s, L, part, x,
                     ps, cs);
  HIPCHK(hipGetLastError());
}

void launch_mg_young(const Lattice &Lf, const Lattice &Lc, const double *yf, double *yc,
                     hipStream_t s, int mode) {
  const dim3 g(blocks_for(Lc.nelx, 256)), t(256);
  switch (mode) {
    case 1: hipLaunchKernelGGL(k_mg_young<1>, g, t, 0, s, Lf, Lc, yf, yc); break;
    case 3: hipLaunchKernelGGL(k_mg_young<3>, g, t, 0, s, Lf, Lc, yf, yc); break;
    case 4: hipLaunchKernelGGL(k_mg_young<4>, g, t, 0, s, Lf, Lc, yf, yc); break;
    default: hipLaunchKernelGGL(k_mg_young<0>, g, t, 0, s, Lf, Lc, yf, yc); break;
  }
  HIPCHK(hipGetLastError());
}

void launch_mg_jacobi(long long n, double omega, const double *dinv, const double *b,
                      const double *y, double *x, hipStream_t s, int nc) {
  if (nc < 1) throw Error(kInternal, "jacobi: no load case");
  hipLaunchKernelGGL(k_mg_jacobi, dim3(blocks_for(n, 256), (unsigned)nc), dim3(256), 0, s, n, omega,
                     dinv, b, y, x);
  HIPCHK(hipGetLastError());
}

unsigned launch_mg_jacobi_dot(long long n, double omega, const double *dinv, const double *b,
                              const double *y, const double *xin, double *xout, const double *q,
                              double *part, hipStream_t s) {
  hipLaunchKernelGGL(k_mg_jacobi_dot, dim3(DOT_BLOCKS), dim3(256), 0, s, n, omega, dinv, b, y, xin,
                     xout, q, part);
  HIPCHK(hipGetLastError());
  return DOT_BLOCKS;
}

void launch_mg_prolong_jacobi(const Lattice &Lf, const Lattice &Lc, const KeCoef &kc,
                              const double *young, const uint8_t *fmask, const double *xcoarse,
                              const double *x, const double *b, const double *dinv, double omega,
                              double *out, hipStream_t s, const double *xcoarse2, int nc,
                              long long sf, long long sc) {
  if (Lf.NX != 2 * Lc.NX - 1 || Lf.NY != 2 * Lc.NY - 1 || Lf.NZ != 2 * Lc.NZ - 1)
    throw Error(kInvalidArgument, "prolongation: lattices are not 2x apart");
  if (out == x) throw Error(kInternal, "prolong_jacobi: out may not alias x");
  if (nc < 1) throw Error(kInternal, "prolong_jacobi: no load case");
  const dim3 g((Lf.NX + QTX - 1) / QTX, (Lf.NY + QTY - 1) / QTY,
               ((Lf.NZ + QTZ - 1) / QTZ) * (unsigned)nc),
      t(256);
  if (kc.sym == 2)
    hipLaunchKernelGGL(k_mg_prolong_jacobi<2>, g, t, 0, s, Lf, Lc, kc.ke, kc.vals, young, fmask,
                       xcoarse, xcoarse2, x, b, dinv, omega, out, sf, sc);
  else if (kc.sym == 1)
    hipLaunchKernelGGL(k_mg_prolong_jacobi<1>, g, t, 0, s, Lf, Lc, kc.ke, kc.vals, young, fmask,
                       xcoarse, xcoarse2, x, b, dinv, omega, out, sf, sc);
  else
    hipLaunchKernelGGL(k_mg_prolong_jacobi<0>, g, t, 0, s, Lf, Lc, kc.ke, kc.vals, young, fmask,
                       xcoarse, xcoarse2, x, b, dinv, omega, out, sf, sc);
  HIPCHK(hipGetLastError());
}

void launch_mg_direct_wres(const Lattice &L, const KeCoef &kc, const double *young,
                           const uint8_t *fmask, const double *xc, const double *b,
                           const double *dinv, double omega, double *u, double *r, double *x0,
                           hipStream_t s, int nc, long long cs) {
  if (xc == u || xc == r || xc == x0) throw Error(kInternal, "direct_wres: xc aliases an output");
  if (nc < 1) throw Error(kInternal, "direct_wres: no load case");
  const dim3 g(blocks_for(L.numnp, 256), (unsigned)nc), t(256);
  if (kc.sym == 2)
    hipLaunchKernelGGL(k_mg_direct_wres<2>, g, t, 0, s, L, kc.ke, kc.vals, young, fmask, xc, b, dinv,
                       omega, u, r, x0, cs);
  else if (kc.sym == 1)
    hipLaunchKernelGGL(k_mg_direct_wres<1>, g, t, 0, s, L, kc.ke, kc.vals, young, fmask, xc, b, dinv,
                       omega, u, r, x0, cs);
  else
    hipLaunchKernelGGL(k_mg_direct_wres<0>, g, t, 0, s, L, kc.ke, kc.vals, young, fmask, xc, b, dinv,
                       omega, u, r, x0, cs);
  HIPCHK(hipGetLastError());
}

void launch_mg_direct(int mode, const Lattice &L, const KeCoef &kc, const double *young,
                      const uint8_t *fmask, const double *x, const double *b, const double *dinv,
                      double omega, double *out, hipStream_t s, int nc, long long cs) {
  if (mode != MG_JACOBI && mode != MG_RESIDUAL) throw Error(kInvalidArgument, "unknown direct mode");
  if (nc < 1) throw Error(kInternal, "direct: no load case");
  if (kc.sym == 2)
    direct_dispatch<2>(mode, L, kc, young, fmask, x, b, dinv, omega, out, s, nc, cs);
  else if (kc.sym == 1)
    direct_dispatch<1>(mode, L, kc, young, fmask, x, b, dinv, omega, out, s, nc, cs);
  else
    direct_dispatch<0>(mode, L, kc, young, fmask, x, b, dinv, omega, out, s, nc, cs);
  HIPCHK(hipGetLastError());
}

}  // namespace icarat_hip
