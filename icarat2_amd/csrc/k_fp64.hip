// k_fp64.hip -- FP64 kernels of the equilibrium solve K(x)u = f on the voxel
// hexa8 lattice: matrix-free K·u, diag(K), fused Jacobi-PCG, residual norms,
// sensitivities and the reference-numbering permutations.
//
// K·u is node-centric: each output node gathers the 27 neighbouring nodal
// displacements from an LDS halo tile and contracts, for each of its 8
// elements, the 3 rows of KE0 that belong to it (3x24 per element) scaled by
// the element's SIMP young.  KE0 lives in constant memory and every index
// into it is a compile-time constant, so its entries reach the FMAs as scalar
// (wave-uniform) operands.  No atomics, no element-result scratch: the result
// is deterministic and each node is written exactly once.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <type_traits>

#include "contract.hpp"
#include "kernels.hpp"
#include "ke_sym_table.hpp"

namespace icarat_hip {

namespace {

constexpr int SX = 32, SY = 8, SZ = 4, NT = 256;
constexpr int HX = SX + 2, HY = SY + 2, HZ = SZ + 2, HN = HX * HY * HZ;
constexpr int EX = SX + 1, EY = SY + 1, EZ = SZ + 1, EN = EX * EY * EZ;
constexpr int VEC_BLOCKS = 2048;
// update kernel of fused batches: fewer blocks, so that every stencil block
// reduces few partials
constexpr int UPD_BLOCKS = 512;

__device__ __forceinline__ bool first_thread() {
  return threadIdx.x == 0 && blockIdx.x == 0 && blockIdx.y == 0 && blockIdx.z == 0;
}

// Beta step of a fused Jacobi-PCG batch, at the start of the stencil kernel of
// in-batch iteration j >= 1: every block reduces the previous update's
// partials (rho_j, r.r) in the same fixed order, so all blocks reach the same
// beta_j = rho_j / rho_{j-1} and the same stop decision; the first thread
// records them.  Returns false when the iteration must not run.
__device__ __forceinline__ bool fused_beta(const double *__restrict__ pu, unsigned Gu, int j,
                                           PcgScalars *S, double *sh, double &beta,
                                           bool leader) {
  double tot[2];
  sum_partials<NT, 2>(pu, Gu, tot, sh);
  const double rho1 = tot[0], rr = tot[1];
  const double rho0 = S->rho_slot[(j - 1) & 1];
  const double rnorm = sqrt(rr);
  const int iter = S->iter_base + j;
  int status = -1;
  if (!isfinite(rnorm) || !isfinite(rho1)) status = 2;
  else if (rnorm < S->tol * sqrt(S->bnorm2) || rnorm == 0.0) status = 0;
  else if (iter > S->maxit) status = 1;
  else if (rnorm < S->repl_norm) status = 3;
  beta = rho1 / rho0;
  if (leader) {
    S->rho = rho1;
    S->rr = rr;
    S->iter = iter;
    S->beta = beta;
    S->rho_slot[j & 1] = rho1;
    if (status >= 0) {
      S->status = status;
      S->done = 1;
    }
  }
  return status < 0;
}

// Alpha step of a fused batch, at the start of the update kernel of in-batch
// iteration j: every block reduces the stencil's p.q partials.
__device__ __forceinline__ bool fused_alpha(const double *__restrict__ ps, unsigned Gs, int j,
                                            PcgScalars *S, double *sh, double &alpha,
                                            bool leader) {
  double tot[1];
  sum_partials<NT, 1>(ps, Gs, tot, sh);
  const double pq = tot[0];
  alpha = S->rho_slot[j & 1] / pq;
  const bool ok = pq > 0.0 && isfinite(alpha);
  if (leader) {
    S->pq = pq;
    S->alpha = alpha;
    if (!ok) {
      S->status = 2;
      S->done = 1;
    }
  }
  return ok;
}

// PCG stencil prologue: beta from the scalars, or (FUSE) from the previous
// update's partials.  Returns false when the kernel must not run.
template <bool PCG, bool FUSE>
__device__ __forceinline__ bool pcg_prologue(PcgScalars *S, const double *pu, unsigned Gu, int j,
                                             double *sh, double &beta,
                                             bool leader = first_thread()) {
  beta = 0.0;
  if constexpr (PCG) {
    if (S->done) return false;
    if constexpr (FUSE) return fused_beta(pu, Gu, j, S, sh, beta, leader);
    beta = S->beta;
  }
  return true;
}

template <bool PCG, int SYM, bool FUSE>
__global__ __launch_bounds__(NT) void k_stencil64(
    Lattice L, const double *__restrict__ c_KE0, KeVals kv, const double *__restrict__ young,
    const uint8_t *__restrict__ fmask, const double *__restrict__ src, const double *__restrict__ dinv,
    const double *__restrict__ p_old, double *__restrict__ p_new, double *__restrict__ y,
    PcgScalars *S, double *partials, const double *__restrict__ pu, unsigned Gu, int j) {
  __shared__ double su[3][HN];
  __shared__ double se[EN];
  __shared__ double sh[8];
  double beta;
  if (!pcg_prologue<PCG, FUSE>(S, pu, Gu, j, sh, beta)) return;
  const int x0 = blockIdx.x * SX, y0 = blockIdx.y * SY, z0 = blockIdx.z * SZ;
  const long long Np = L.Np;
  for (int t = threadIdx.x; t < HN; t += NT) {
    const int hx = t % HX, hy = (t / HX) % HY, hz = t / (HX * HY);
    const int gx = x0 - 1 + hx, gy = y0 - 1 + hy, gz = z0 - 1 + hz;
    double v0 = 0.0, v1 = 0.0, v2 = 0.0;
    if (gx >= 0 && gx < L.NX && gy >= 0 && gy < L.NY && gz >= 0 && gz < L.NZ) {
      const long long n = gx + (long long)L.NX * (gy + (long long)L.NY * gz);
      if constexpr (PCG) {
        // dinv == nullptr: p = src + beta p_old (src is already preconditioned)
        v0 = (dinv ? dinv[n] * src[n] : src[n]) + beta * p_old[n];
        v1 = (dinv ? dinv[Np + n] * src[Np + n] : src[Np + n]) + beta * p_old[Np + n];
        v2 = (dinv ? dinv[2 * Np + n] * src[2 * Np + n] : src[2 * Np + n]) + beta * p_old[2 * Np + n];
        if (hx >= 1 && hx <= SX && hy >= 1 && hy <= SY && hz >= 1 && hz <= SZ) {
          p_new[n] = v0;
          p_new[Np + n] = v1;
          p_new[2 * Np + n] = v2;
        }
      } else {
        v0 = src[n];
        v1 = src[Np + n];
        v2 = src[2 * Np + n];
      }
    }
    su[0][t] = v0;
    su[1][t] = v1;
    su[2][t] = v2;
  }
  for (int t = threadIdx.x; t < EN; t += NT) {
    const int ex = t % EX, ey = (t / EX) % EY, ez = t / (EX * EY);
    const int gx = x0 - 1 + ex, gy = y0 - 1 + ey, gz = z0 - 1 + ez;
    double v = 0.0;
    if (gx >= 0 && gx < L.nx && gy >= 0 && gy < L.ny && gz >= 0 && gz < L.nz)
      v = young[gx + (long long)L.nx * (gy + (long long)L.ny * gz)];
    se[t] = v;
  }
  __syncthreads();

  const int tx = threadIdx.x % SX, ty = threadIdx.x / SX;
  double pq = 0.0;
  for (int iz = 0; iz < SZ; ++iz) {
    const int gx = x0 + tx, gy = y0 + ty, gz = z0 + iz;
    if (gx < L.NX && gy < L.NY && gz < L.NZ) {
      double U[27][3], E[8];
#pragma unroll
      for (int dz = -1; dz <= 1; ++dz)
#pragma unroll
        for (int dy = -1; dy <= 1; ++dy)
#pragma unroll
          for (int dx = -1; dx <= 1; ++dx) {
            const int h = (tx + 1 + dx) + HX * ((ty + 1 + dy) + HY * (iz + 1 + dz));
#pragma unroll
            for (int c = 0; c < 3; ++c) U[nb(dx, dy, dz)][c] = su[c][h];
          }
#pragma unroll
      for (int s = 0; s < 8; ++s) {
        const int ex = (s & 1) - 1, ey = ((s >> 1) & 1) - 1, ez = ((s >> 2) & 1) - 1;
        E[s] = se[(tx + 1 + ex) + EX * ((ty + 1 + ey) + EY * (iz + 1 + ez))];
      }
      double out[3];
      node_contract<SYM>(c_KE0, kv, U, E, out);
      const long long n = gx + (long long)L.NX * (gy + (long long)L.NY * gz);
      const unsigned m = fmask[n];
#pragma unroll
      for (int c = 0; c < 3; ++c) out[c] = ((m >> c) & 1u) ? out[c] : 0.0;
      y[n] = out[0];
      y[Np + n] = out[1];
      y[2 * Np + n] = out[2];
      if constexpr (PCG) pq += U[13][0] * out[0] + U[13][1] * out[1] + U[13][2] * out[2];
    }
  }
  if constexpr (PCG) {
    double v[1] = {pq};
    store_partials<NT, 1>(v, partials, sh);
  }
}

// ---- z-marching variant ----------------------------------------------------
// A workgroup owns a ZX x ZY column of nodes and marches over the element
// layers in z.  Each thread keeps the 3x3 node neighbourhood of two planes
// (the layer's bottom and top) in registers; the layer's 4 elements around
// the thread's node column give the last contribution of the bottom node
// (finished and written) and the first of the top node (kept in registers).
// The next plane and element layer are fetched from global memory before the
// layer is contracted and reach a double-buffered LDS ring afterwards.
constexpr int ZX = 32, ZY = 8;
constexpr int ZPX = ZX + 2, ZPY = ZY + 2, ZPN = ZPX * ZPY;
constexpr int ZLX = ZX + 1, ZLY = ZY + 1, ZLN = ZLX * ZLY;
constexpr int ZPER = (ZPN + NT - 1) / NT;  // plane nodes per thread
constexpr int ZEPER = (ZLN + NT - 1) / NT; // layer elements per thread
// output of the z-marching K.u: the node-value type, FP64 for the fused sweep
template <int EPI, class T>
using OutT = typename std::conditional<EPI != 0, double, T>::type;

// T is the type of the node values, LDS planes and contraction (double; float
// for the multigrid smoother's FP32 fine level, PCG = false only).
template <bool PCG, class T = double>
struct PlaneFetch {
  T v[ZPER][3];
  double d[ZPER][3], pold[ZPER][3];  // PCG: dinv and p_old (v holds r)
  T e[ZEPER];
};

template <bool PCG, class T>
__device__ __forceinline__ void fetch_plane(const Lattice &L, int x0, int y0, int zp,
                                            const T *__restrict__ src,
                                            const double *__restrict__ dinv,
                                            const double *__restrict__ p_old,
                                            const T *__restrict__ young, PlaneFetch<PCG, T> &F) {
#pragma unroll
  for (int k = 0; k < ZPER; ++k) {
    const int t = threadIdx.x + k * NT;
    const int gx = x0 - 1 + t % ZPX, gy = y0 - 1 + t / ZPX;
    const bool in = t < ZPN && zp >= 0 && zp < L.NZ && gx >= 0 && gx < L.NX && gy >= 0 && gy < L.NY;
    const long long n = in ? gx + (long long)L.NX * (gy + (long long)L.NY * zp) : 0;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      F.v[k][c] = in ? src[c * L.Np + n] : T(0);
      if constexpr (PCG) {
        // dinv == nullptr: p = src + beta p_old (src is already preconditioned)
        F.d[k][c] = in ? (dinv ? dinv[c * L.Np + n] : 1.0) : 0.0;
        F.pold[k][c] = in ? p_old[c * L.Np + n] : 0.0;
      }
    }
  }
  // element layer zp-1 (between node planes zp-1 and zp)
  const int ze = zp - 1;
#pragma unroll
  for (int k = 0; k < ZEPER; ++k) {
    const int t = threadIdx.x + k * NT;
    const int ex = x0 - 1 + t % ZLX, ey = y0 - 1 + t / ZLX;
    const bool in = t < ZLN && ze >= 0 && ze < L.nz && ex >= 0 && ex < L.nx && ey >= 0 && ey < L.ny;
    const long long ei = in ? ex + (long long)L.nx * (ey + (long long)L.ny * ze) : 0;
    F.e[k] = in ? young[ei] : T(0);
  }
}

// Writes a fetched plane (and layer) into LDS; PCG: forms p = dinv r + beta
// p_old and stores it to p_new for owned nodes (own_x0: the halo column at
// x = 0 is owned too, for tiles that start at x = 1).
template <bool PCG, class T>
__device__ __forceinline__ void commit_plane(const Lattice &L, int x0, int y0, int zp, bool own_z,
                                             bool own_x0, double beta, const PlaneFetch<PCG, T> &F,
                                             double *__restrict__ p_new, T (*sp)[ZPN], T *se) {
#pragma unroll
  for (int k = 0; k < ZPER; ++k) {
    const int t = threadIdx.x + k * NT;
    if (t >= ZPN) continue;
    T v[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      if constexpr (PCG) v[c] = F.d[k][c] * F.v[k][c] + beta * F.pold[k][c];
      else v[c] = F.v[k][c];
      sp[c][t] = v[c];
    }
    if constexpr (PCG) {
      const int hx = t % ZPX, hy = t / ZPX;
      const int gx = x0 - 1 + hx, gy = y0 - 1 + hy;
      const bool own_col = (hx >= 1 && hx <= ZX) || (own_x0 && hx == 0);
      if (own_z && own_col && hy >= 1 && hy <= ZY && gx < L.NX && gy < L.NY) {
        const long long n = gx + (long long)L.NX * (gy + (long long)L.NY * zp);
#pragma unroll
        for (int c = 0; c < 3; ++c) p_new[c * L.Np + n] = v[c];
      }
    }
  }
#pragma unroll
  for (int k = 0; k < ZEPER; ++k) {
    const int t = threadIdx.x + k * NT;
    if (t < ZLN) se[t] = F.e[k];
  }
}

// 3x3 neighbourhood of the thread's column in one LDS plane
template <class T>
__device__ __forceinline__ void read_nbhd(const T (*sp)[ZPN], int tx, int ty, T (&U)[9][3]) {
#pragma unroll
  for (int dy = -1; dy <= 1; ++dy)
#pragma unroll
    for (int dx = -1; dx <= 1; ++dx) {
      const int h = (tx + 1 + dx) + ZPX * (ty + 1 + dy);
#pragma unroll
      for (int c = 0; c < 3; ++c) U[(dx + 1) + 3 * (dy + 1)][c] = sp[c][h];
    }
}

// Contributions of the 4 elements of one layer around the node column:
// bot = rows of the bottom node (local z = 0), top = rows of the top node.
template <int SYM, class T>
__device__ __forceinline__ void layer_contract(const T *__restrict__ c_KE0, const KeValsT<T> &kv,
                                               const T (&Ulo)[9][3], const T (&Uhi)[9][3],
                                               const T (&E)[4], T (&bot)[3], T (&top)[3]) {
  bot[0] = bot[1] = bot[2] = T(0);
  top[0] = top[1] = top[2] = T(0);
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    const int ex = (s & 1) - 1, ey = (s >> 1) - 1;
    const int ab = lnode(-ex, -ey, 0), at = lnode(-ex, -ey, 1);
    T tb0 = T(0), tb1 = T(0), tb2 = T(0), tt0 = T(0), tt1 = T(0), tt2 = T(0);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const int o = (ex + lnode_x(b) + 1) + 3 * (ey + lnode_y(b) + 1);
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        const T u = lnode_z(b) ? Uhi[o][j] : Ulo[o][j];
        tb0 += kcoef<SYM>(c_KE0, kv, (3 * ab + 0) * 24 + 3 * b + j) * u;
        tb1 += kcoef<SYM>(c_KE0, kv, (3 * ab + 1) * 24 + 3 * b + j) * u;
        tb2 += kcoef<SYM>(c_KE0, kv, (3 * ab + 2) * 24 + 3 * b + j) * u;
        tt0 += kcoef<SYM>(c_KE0, kv, (3 * at + 0) * 24 + 3 * b + j) * u;
        tt1 += kcoef<SYM>(c_KE0, kv, (3 * at + 1) * 24 + 3 * b + j) * u;
        tt2 += kcoef<SYM>(c_KE0, kv, (3 * at + 2) * 24 + 3 * b + j) * u;
      }
    }
    bot[0] += E[s] * tb0;
    bot[1] += E[s] * tb1;
    bot[2] += E[s] * tb2;
    top[0] += E[s] * tt0;
    top[1] += E[s] * tt1;
    top[2] += E[s] * tt2;
  }
}

// Epilogues of the z-marching K.u (EPI, with PCG = false): 0 y = K x.  1 the
// multigrid smoother's last fine sweep: the node's own value x is the centre
// of the bottom plane, so y receives x' = x + omega dinv (b - K x) (out of
// place: neighbouring columns and the next z-chunk still read x) and the block
// partials of x'.b and x'.q.  2 the same sweep without partials (coarser
// z-marching levels), or the start of a W-cycle's second visit (jac_epilogue).
// The epilogues compute in FP64 and y, b, dinv, q are FP64 whatever T is.

// The work split of a z-marching launch.  Tiles of ZX x ZY node columns start
// at node column x = xlo: 1 when column x = 0 holds no free DOF (a clamped
// face; its K.u is 0 and the first tile column writes it from its halo), else
// 0.  The W = tiles x NZ tile-planes are dealt in equal consecutive ranges
// (tile-major, z fastest) to the nb blocks, so one round of resident
// workgroups covers the lattice however its tiles divide it; a block marches
// one segment per tile its range touches.
struct ZSplit {
  int xlo, nbx, nby, nb;
  long long W;
};

// XCD-aware tile order: workgroups go to the 8 XCDs round-robin by id, so
// each XCD is given a contiguous range of the nt work items (columns fastest):
// one XCD marches neighbouring columns of one z-slab and shares their halos in
// its L2.  xcd_order maps a workgroup id to its work item, xcd_slot back.
__device__ __forceinline__ unsigned xcd_order(unsigned b, unsigned nt) {
  const unsigned q = nt / 8, rem = nt % 8, xcd = b % 8, slot = b / 8;
  return (xcd < rem ? xcd * (q + 1) : rem * (q + 1) + (xcd - rem) * q) + slot;
}
__device__ __forceinline__ unsigned xcd_slot(unsigned t, unsigned nt) {
  const unsigned q = nt / 8, rem = nt % 8, big = rem * (q + 1);
  if (t < big) return (t % (q + 1)) * 8 + t / (q + 1);
  return ((t - big) % q) * 8 + rem + (t - big) / q;
}

// The epilogues of EPI 2 at an output node n with centre value xc and masked
// K x = kx: the smoother sweep y = xc + omega dinv (b - kx) (no partials); or,
// with je.u, u = xc, y = b - kx and x0 = omega dinv y.  In FP64.  (EPI 1, the
// fine level's last sweep with the y.b and y.q partials, is written out at
// its call sites.)
template <class T, class TO>
__device__ __forceinline__ void jac_epilogue(const JacEpilogue &je, long long Np, long long n,
                                             const T (&xc)[3], const T (&kx)[3],
                                             const double *__restrict__ dinv, TO *__restrict__ y) {
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const long long ic = c * Np + n;
    const double bc = je.b[ic];
    if (je.u) {
      const double r = bc - (double)kx[c];
      je.u[ic] = (double)xc[c];
      y[ic] = r;
      if (je.x0) je.x0[ic] = je.omega * dinv[ic] * r;
    } else {
      y[ic] = (double)xc[c] + je.omega * dinv[ic] * (bc - (double)kx[c]);
    }
  }
}

// One segment of a block's z-marching K.u: node planes [z0, z1) of tile t
// (element layers z0-1 .. z1-1: the layer below z0 completes plane z0).  The
// block partials accumulate into pq (PCG) and zr, zq (EPI 1).
template <bool PCG, int SYM, int EPI, class T, class TO>
__device__ __forceinline__ void stencil64z_seg(
    const Lattice &L, const ZSplit &Z, int t, int z0, int z1, double beta,
    const T *__restrict__ c_KE0, const KeValsT<T> &kv, const T *__restrict__ young,
    const uint8_t *__restrict__ fmask, const T *__restrict__ src, const double *__restrict__ dinv,
    const double *__restrict__ p_old, double *__restrict__ p_new, TO *__restrict__ y,
    const JacEpilogue &je, T (*sp)[3][ZPN], T (*se)[ZLN], double &pq, double &zr, double &zq) {
  const int x0 = Z.xlo + (t % Z.nbx) * ZX, y0 = (t / Z.nbx) * ZY;
  const int tx = threadIdx.x % ZX, ty = threadIdx.x / ZX;
  const int gx = x0 + tx, gy = y0 + ty;
  const bool col_in = gx < L.NX && gy < L.NY;
  // tiles starting at x = 1: the first tile column also writes node column
  // x = 0 (no free DOF, K.u = 0 there) from the halo it holds
  const bool own_x0 = Z.xlo == 1 && x0 == 1;
  const bool edge = own_x0 && tx == 0 && gy < L.NY;
  PlaneFetch<PCG, T> F;
  T Ulo[9][3], Uhi[9][3];
  // Slot s holds node plane zb+1 and element layer zb of the step that reads
  // it.  Plane z0-1 goes to registers through slot 1.
  fetch_plane<PCG, T>(L, x0, y0, z0 - 1, src, dinv, p_old, young, F);
  commit_plane<PCG, T>(L, x0, y0, z0 - 1, false, own_x0, beta, F, p_new, sp[1], se[1]);
  __syncthreads();
  read_nbhd(sp[1], tx, ty, Ulo);
  __syncthreads();
  // plane z0 and layer z0-1 -> slot 0 for step 0
  fetch_plane<PCG, T>(L, x0, y0, z0, src, dinv, p_old, young, F);
  commit_plane<PCG, T>(L, x0, y0, z0, z0 < z1, own_x0, beta, F, p_new, sp[0], se[0]);
  __syncthreads();
  T acc[3] = {T(0), T(0), T(0)};
  const int nsteps = z1 - z0 + 1;  // element layers z0-1 .. z1-1
  for (int k = 0; k < nsteps; ++k) {
    const int zb = z0 - 1 + k;
    const int cur = k & 1, nxt = cur ^ 1;
    const bool more = k + 1 < nsteps;
    // prefetch plane zb+2 with layer zb+1 for the next step
    if (more) fetch_plane<PCG, T>(L, x0, y0, zb + 2, src, dinv, p_old, young, F);
    read_nbhd(sp[cur], tx, ty, Uhi);
    T E[4];
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      const int ex = (s & 1) - 1, ey = (s >> 1) - 1;
      E[s] = se[cur][(tx + 1 + ex) + ZLX * (ty + 1 + ey)];
    }
    T bot[3], top[3];
    layer_contract<SYM>(c_KE0, kv, Ulo, Uhi, E, bot, top);
    if (k >= 1 && col_in) {
      const long long n = gx + (long long)L.NX * (gy + (long long)L.NY * zb);
      const unsigned m = fmask[n];
      T out[3];
#pragma unroll
      for (int c = 0; c < 3; ++c) out[c] = ((m >> c) & 1u) ? acc[c] + bot[c] : T(0);
      if constexpr (EPI == 1) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          const long long ic = c * L.Np + n;
          const double bc = je.b[ic];
          const double xn = (double)Ulo[4][c] + je.omega * dinv[ic] * (bc - (double)out[c]);
          y[ic] = xn;
          zr += xn * bc;
          zq += xn * je.q[ic];
        }
      } else if constexpr (EPI == 2) {
        jac_epilogue(je, L.Np, n, Ulo[4], out, dinv, y);
      } else {
#pragma unroll
        for (int c = 0; c < 3; ++c) y[c * L.Np + n] = (TO)out[c];
      }
      if constexpr (PCG) pq += Ulo[4][0] * out[0] + Ulo[4][1] * out[1] + Ulo[4][2] * out[2];
    }
    if (k >= 1 && edge) {
      // node (0, gy, zb): every component constrained, K.u masked to 0
      const long long n = (long long)L.NX * (gy + (long long)L.NY * zb);
      if constexpr (EPI == 1) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          const long long ic = c * L.Np + n;
          const double bc = je.b[ic];
          const double xn = (double)Ulo[3][c] + je.omega * dinv[ic] * (bc - (double)T(0));
          y[ic] = xn;
          zr += xn * bc;
          zq += xn * je.q[ic];
        }
      } else if constexpr (EPI == 2) {
        const T zero[3] = {T(0), T(0), T(0)};
        jac_epilogue(je, L.Np, n, Ulo[3], zero, dinv, y);
      } else {
#pragma unroll
        for (int c = 0; c < 3; ++c) y[c * L.Np + n] = (TO)T(0);
      }
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) acc[c] = top[c];
#pragma unroll
    for (int o = 0; o < 9; ++o)
#pragma unroll
      for (int c = 0; c < 3; ++c) Ulo[o][c] = Uhi[o][c];
    if (more)
      commit_plane<PCG, T>(L, x0, y0, zb + 2, zb + 2 < z1, own_x0, beta, F, p_new, sp[nxt],
                           se[nxt]);
    __syncthreads();
  }
}

// A block's share of a z-marching K.u: work range bw of the split Z (its
// consecutive tile-planes, one segment per tile they touch); the block's
// partials go to slot pslot of Gp.
template <bool PCG, int SYM, bool FUSE, int EPI, class T, class TO>
__device__ __forceinline__ void stencil64z_work(
    const Lattice &L, const ZSplit &Z, unsigned bw, const T *__restrict__ c_KE0,
    const KeValsT<T> &kv, const T *__restrict__ young, const uint8_t *__restrict__ fmask,
    const T *__restrict__ src, const double *__restrict__ dinv, const double *__restrict__ p_old,
    double *__restrict__ p_new, TO *__restrict__ y, PcgScalars *S, double *partials,
    const double *__restrict__ pu, unsigned Gu, int j, const JacEpilogue &je, unsigned Gp,
    unsigned pslot, bool leader, T (*sp)[3][ZPN], T (*se)[ZLN], double *sh) {
  static_assert(!PCG || std::is_same<T, double>::value, "the PCG stencil is FP64");
  double beta;
  if (!pcg_prologue<PCG, FUSE>(S, pu, Gu, j, sh, beta, leader)) return;
  double pq = 0.0, zr = 0.0, zq = 0.0;
  const long long u1 = (long long)(bw + 1) * Z.W / Z.nb;
  for (long long u = (long long)bw * Z.W / Z.nb; u < u1;) {
    const int t = (int)(u / L.NZ);
    const int z0 = (int)(u - (long long)t * L.NZ);
    const int z1 = (int)std::min<long long>(L.NZ, u1 - (long long)t * L.NZ);
    stencil64z_seg<PCG, SYM, EPI, T, TO>(L, Z, t, z0, z1, beta, c_KE0, kv, young, fmask, src, dinv,
                                         p_old, p_new, y, je, sp, se, pq, zr, zq);
    u = (long long)t * L.NZ + z1;
  }
  if constexpr (PCG) {
    double v[1] = {pq};
    store_partials_at<NT, 1>(v, partials, sh, Gp, pslot);
  }
  if constexpr (EPI == 1) {
    double v[2] = {zr, zq};
    store_partials_at<NT, 2>(v, partials, sh, Gp, pslot);
  }
}

template <bool PCG, int SYM, bool FUSE, int EPI, class T = double, class TO = OutT<EPI, T>>
__global__ __launch_bounds__(NT) void k_stencil64z(
    Lattice L, ZSplit Z, const T *__restrict__ c_KE0, KeValsT<T> kv, const T *__restrict__ young,
    const uint8_t *__restrict__ fmask, const T *__restrict__ src,
    const double *__restrict__ dinv, const double *__restrict__ p_old, double *__restrict__ p_new,
    TO *__restrict__ y, PcgScalars *S, double *partials, const double *__restrict__ pu,
    unsigned Gu, int j, JacEpilogue je) {
  __shared__ T sp[2][3][ZPN];
  __shared__ T se[2][ZLN];
  __shared__ double sh[8];
  const unsigned b = blockIdx.x;
  stencil64z_work<PCG, SYM, FUSE, EPI, T, TO>(L, Z, xcd_order(b, Z.nb), c_KE0, kv, young, fmask,
                                              src, dinv, p_old, p_new, y, S, partials, pu, Gu, j,
                                              je, Z.nb, b, first_thread(), sp, se, sh);
}

// Several right-hand sides with one operator (load cases): nc cases of a
// Jacobi-PCG batch step in one launch.  Case c's vectors start at c * cs
// (p: two buffers per case, c * 2 cs, chosen by the parity of the case's
// iteration count at the batch start, as the single-case batches do); its
// scalars are S[c]; its partials are stored at [c * nb, (c+1) * nb) in the
// slots of the single-case launch (same split), so each case's sums equal
// those of its own solve bit for bit.  Work items run case-fastest: the cases
// of one work range land on the same XCD together and read its young and
// mask through its L2.
template <int SYM, bool FUSE>
__global__ __launch_bounds__(NT) void k_stencil64z_mc(
    Lattice L, ZSplit Z, const double *__restrict__ c_KE0, KeVals kv,
    const double *__restrict__ young, const uint8_t *__restrict__ fmask,
    const double *__restrict__ r, const double *__restrict__ dinv, double *__restrict__ p,
    double *__restrict__ q, PcgScalars *S, double *partials, const double *__restrict__ pu,
    unsigned Gu, int j, int nc, long long cs) {
  __shared__ double sp[2][3][ZPN];
  __shared__ double se[2][ZLN];
  __shared__ double sh[8];
  const unsigned ntot = gridDim.x, nt = (unsigned)Z.nb;
  const unsigned w = xcd_order(blockIdx.x, ntot);
  const int c = (int)(w % nc);
  const unsigned T_ = w / nc;
  PcgScalars *Sc = S + c;
  const int par = (j + Sc->iter_base) & 1;
  double *pc = p + 2 * cs * c;
  stencil64z_work<true, SYM, FUSE, false, double, double>(
      L, Z, T_, c_KE0, kv, young, fmask, r + cs * c, dinv, pc + cs * par, pc + cs * (par ^ 1),
      q + cs * c, Sc, partials + (size_t)nt * c, pu + (size_t)2 * Gu * c, Gu, j, JacEpilogue{},
      nt, xcd_slot(T_, nt), T_ == 0 && threadIdx.x == 0, sp, se, sh);
}

// The multigrid solve's PCG stencil for nc load cases in one launch: case c
// forms p_new = z + beta_c p_old from its V-cycle result z (no D^-1), q = K p
// and its p.q partials, each at c x cs; its scalars are S[c], its partials at
// [c x nb, (c+1) x nb) in the slots of the single-case launch, so every case's
// sums equal those of its own solve.  Work items run case-fastest.
template <int SYM>
__global__ __launch_bounds__(NT) void k_stencil64z_mcz(
    Lattice L, ZSplit Z, const double *__restrict__ c_KE0, KeVals kv,
    const double *__restrict__ young, const uint8_t *__restrict__ fmask,
    const double *__restrict__ z, const double *__restrict__ p_old, double *__restrict__ p_new,
    double *__restrict__ q, PcgScalars *S, double *partials, int nc, long long cs) {
  __shared__ double sp[2][3][ZPN];
  __shared__ double se[2][ZLN];
  __shared__ double sh[8];
  const unsigned ntot = gridDim.x, nt = (unsigned)Z.nb;
  const unsigned w = xcd_order(blockIdx.x, ntot);
  const int c = (int)(w % nc);
  const unsigned T_ = w / nc;
  stencil64z_work<true, SYM, false, 0, double, double>(
      L, Z, T_, c_KE0, kv, young, fmask, z + cs * c, nullptr, p_old + cs * c, p_new + cs * c,
      q + cs * c, S + c, partials + (size_t)nt * c, nullptr, 0u, 0, JacEpilogue{}, nt,
      xcd_slot(T_, nt), T_ == 0 && threadIdx.x == 0, sp, se, sh);
}

// K.u (EPI 0), the fine level's last sweep with its partials (EPI 1) or an
// EPI 2 epilogue (a coarser level's sweep, or the W-cycle's second-visit
// start) for nc load cases in one launch: case c's x, y, b, q, u, x0 at c x cs
// (D^-1 shared), its two partials per block (EPI 1) at [2 c nb, 2 (c+1) nb) in
// the layout of the single-case launch.
template <int SYM, int EPI>
__global__ __launch_bounds__(NT) void k_stencil64z_mcv(
    Lattice L, ZSplit Z, const double *__restrict__ c_KE0, KeVals kv,
    const double *__restrict__ young, const uint8_t *__restrict__ fmask,
    const double *__restrict__ x, const double *__restrict__ dinv, double *__restrict__ y,
    double *partials, JacEpilogue je, int nc, long long cs) {
  __shared__ double sp[2][3][ZPN];
  __shared__ double se[2][ZLN];
  __shared__ double sh[8];
  const unsigned ntot = gridDim.x, nt = (unsigned)Z.nb;
  const unsigned w = xcd_order(blockIdx.x, ntot);
  const int c = (int)(w % nc);
  const unsigned T_ = w / nc;
  JacEpilogue jc = je;
  if (jc.b) jc.b += cs * c;
  if (jc.q) jc.q += cs * c;
  if (jc.u) jc.u += cs * c;
  if (jc.x0) jc.x0 += cs * c;
  stencil64z_work<false, SYM, false, EPI, double, double>(
      L, Z, T_, c_KE0, kv, young, fmask, x + cs * c, dinv, nullptr, nullptr, y + cs * c, nullptr,
      EPI == 1 ? partials + (size_t)2 * nt * c : nullptr, nullptr, 0u, 0, jc, nt, xcd_slot(T_, nt),
      false, sp, se, sh);
}

__global__ __launch_bounds__(NT) void k_pcg_update64(long long n, double *__restrict__ x,
                                                     double *__restrict__ r,
                                                     const double *__restrict__ p,
                                                     const double *__restrict__ q,
                                                     const double *__restrict__ dinv,
                                                     PcgScalars *S, double *partials,
                                                     unsigned *counter) {
  __shared__ double sh[8];
  if (S->done) return;
  const double alpha = S->alpha;
  double rz = 0.0, rr = 0.0;
  const long long stride = (long long)gridDim.x * NT;
  for (long long i = blockIdx.x * (long long)NT + threadIdx.x; i < n; i += stride) {
    x[i] += alpha * p[i];
    const double ri = r[i] - alpha * q[i];
    r[i] = ri;
    rz += (dinv[i] * ri) * ri;
    rr += ri * ri;
  }
  double v[2] = {rz, rr};
  store_partials<NT, 2>(v, partials, sh);
}

// Vector half of a fused batch: alpha from the stencil's partials in every
// block, then x += alpha p, r -= alpha q and the partials of r.D^-1r, r.r.
__global__ __launch_bounds__(NT) void k_pcg_update64f(long long n, double *__restrict__ x,
                                                      double *__restrict__ r,
                                                      const double *__restrict__ p,
                                                      const double *__restrict__ q,
                                                      const double *__restrict__ dinv,
                                                      PcgScalars *S, double *partials,
                                                      const double *__restrict__ ps, unsigned Gs,
                                                      int j) {
  __shared__ double sh[8];
  if (S->done) return;
  double alpha;
  if (!fused_alpha(ps, Gs, j, S, sh, alpha, first_thread())) return;
  double rz = 0.0, rr = 0.0;
  const long long stride = (long long)gridDim.x * NT;
  for (long long i = blockIdx.x * (long long)NT + threadIdx.x; i < n; i += stride) {
    x[i] += alpha * p[i];
    const double ri = r[i] - alpha * q[i];
    r[i] = ri;
    rz += (dinv[i] * ri) * ri;
    rr += ri * ri;
  }
  double v[2] = {rz, rr};
  store_partials<NT, 2>(v, partials, sh);
}

// k_pcg_update64f for nc load cases (blockIdx.y = case, the layout of
// k_stencil64z_mc); each case's partials equal those of its own solve.
__global__ __launch_bounds__(NT) void k_pcg_update64f_mc(long long n, double *__restrict__ x,
                                                         double *__restrict__ r,
                                                         const double *__restrict__ p,
                                                         const double *__restrict__ q,
                                                         const double *__restrict__ dinv,
                                                         PcgScalars *S, double *partials,
                                                         const double *__restrict__ ps,
                                                         unsigned Gs, int j, long long cs) {
  __shared__ double sh[8];
  const int c = blockIdx.y;
  PcgScalars *Sc = S + c;
  if (Sc->done) return;
  double alpha;
  if (!fused_alpha(ps + (size_t)Gs * c, Gs, j, Sc, sh, alpha, blockIdx.x == 0 && threadIdx.x == 0))
    return;
  const int par = (j + Sc->iter_base + 1) & 1;
  double *xc = x + cs * c, *rc = r + cs * c;
  const double *pc = p + 2 * cs * c + cs * par, *qc = q + cs * c;
  double rz = 0.0, rr = 0.0;
  const long long stride = (long long)gridDim.x * NT;
  for (long long i = blockIdx.x * (long long)NT + threadIdx.x; i < n; i += stride) {
    xc[i] += alpha * pc[i];
    const double ri = rc[i] - alpha * qc[i];
    rc[i] = ri;
    rz += (dinv[i] * ri) * ri;
    rr += ri * ri;
  }
  double v[2] = {rz, rr};
  store_partials_at<NT, 2>(v, partials + (size_t)2 * gridDim.x * c, sh, gridDim.x, blockIdx.x);
}

__global__ __launch_bounds__(NT) void k_pcg_init64(long long n, const double *__restrict__ b,
                                                   double *__restrict__ x,
                                                   double *__restrict__ r,
                                                   double *__restrict__ p0,
                                                   const double *__restrict__ dinv,
                                                   PcgScalars *S, double *partials,
                                                   unsigned *counter) {
  __shared__ double sh[8];
  double rz = 0.0, rr = 0.0;
  const long long stride = (long long)gridDim.x * NT;
  for (long long i = blockIdx.x * (long long)NT + threadIdx.x; i < n; i += stride) {
    const double bi = b[i];
    x[i] = 0.0;
    r[i] = bi;
    p0[i] = 0.0;
    rz += (dinv[i] * bi) * bi;
    rr += bi * bi;
  }
  double v[2] = {rz, rr};
  store_partials<NT, 2>(v, partials, sh);
}

// r = b - y (true residual after K x), rho = r.dinv.r, rr = r.r
__global__ __launch_bounds__(NT) void k_replace64(long long n, const double *__restrict__ b,
                                                  const double *__restrict__ y,
                                                  double *__restrict__ r,
                                                  const double *__restrict__ dinv,
                                                  double *partials) {
  __shared__ double sh[8];
  double rz = 0.0, rr = 0.0;
  const long long stride = (long long)gridDim.x * NT;
  for (long long i = blockIdx.x * (long long)NT + threadIdx.x; i < n; i += stride) {
    const double ri = b[i] - y[i];
    r[i] = ri;
    rz += (dinv[i] * ri) * ri;
    rr += ri * ri;
  }
  double v[2] = {rz, rr};
  store_partials<NT, 2>(v, partials, sh);
}

__global__ __launch_bounds__(NT) void k_residual_norms(long long n, const double *__restrict__ b,
                                                       const double *__restrict__ y, double *out2,
                                                       double *partials, unsigned *counter) {
  __shared__ double sh[8];
  double a = 0.0, c = 0.0;
  const long long stride = (long long)gridDim.x * NT;
  for (long long i = blockIdx.x * (long long)NT + threadIdx.x; i < n; i += stride) {
    const double d = b[i] - y[i];
    a += d * d;
    c += b[i] * b[i];
  }
  double v[2] = {a, c};
  store_partials<NT, 2>(v, partials, sh);
}

__global__ __launch_bounds__(NT) void k_dot(long long n, const double *__restrict__ a,
                                            const double *__restrict__ b, double *out,
                                            double *partials, unsigned *counter) {
  __shared__ double sh[4];
  double s = 0.0;
  const long long stride = (long long)gridDim.x * NT;
  for (long long i = blockIdx.x * (long long)NT + threadIdx.x; i < n; i += stride) s += a[i] * b[i];
  double v[1] = {s};
  store_partials<NT, 1>(v, partials, sh);
}

__global__ void k_young(const double *__restrict__ x, double *__restrict__ young, long long n,
                        double y1, double y2, double p) {
  // no contraction: the same two roundings as the host's SIMP law
#pragma clang fp contract(off)
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double xp = simp_pow(x[i], p);
  young[i] = (1.0 - xp) * y1 + xp * y2;
}

__global__ void k_diag_inv(Lattice L, const double *__restrict__ c_KE0,
                           const double *__restrict__ young,
                           const uint8_t *__restrict__ fmask, double *__restrict__ dinv,
                           double *__restrict__ diag) {
  const long long n = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (n >= L.numnp) return;
  const int i = (int)(n % L.NX), j = (int)((n / L.NX) % L.NY), k = (int)(n / ((long long)L.NX * L.NY));
  double d[3] = {0.0, 0.0, 0.0};
#pragma unroll
  for (int s = 0; s < 8; ++s) {
    const int ex = (s & 1) - 1, ey = ((s >> 1) & 1) - 1, ez = ((s >> 2) & 1) - 1;
    const int a = i + ex, b = j + ey, c = k + ez;
    if (a < 0 || b < 0 || c < 0 || a >= L.nx || b >= L.ny || c >= L.nz) continue;
    const double E = young[a + (long long)L.nx * (b + (long long)L.ny * c)];
    const int la = lnode(-ex, -ey, -ez);
#pragma unroll
    for (int q = 0; q < 3; ++q) d[q] += E * c_KE0[(3 * la + q) * 24 + 3 * la + q];
  }
  const unsigned m = fmask[n];
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    const bool f = (m >> q) & 1u;
    dinv[q * L.Np + n] = f ? 1.0 / d[q] : 0.0;
    if (diag) diag[q * L.Np + n] = f ? d[q] : 0.0;
  }
}

__global__ void k_flush_small(long long n, const double *__restrict__ u, double *__restrict__ v) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double a = u[i];
  v[i] = fabs(a) < 1.0e-15 ? 0.0 : a;
}

// One thread per element; u_e gathered straight from the nodal arrays (each
// node is shared by 8 neighbouring threads, so the gathers hit L1/L2).
__global__ __launch_bounds__(NT) void k_sensitivity64(Lattice L, const double *__restrict__ c_KE0,
                                                      const double *__restrict__ dens,
                                                      const double *__restrict__ nv, double y1,
                                                      double y2, double p, double vol_e,
                                                      double vnorm, double *__restrict__ dfds,
                                                      double *__restrict__ dgds) {
  const long long e = blockIdx.x * (long long)NT + threadIdx.x;
  if (e >= L.nelx) return;
  const int i = (int)(e % L.nx), j = (int)((e / L.nx) % L.ny), k = (int)(e / ((long long)L.nx * L.ny));
  const long long n0 = i + (long long)L.NX * (j + (long long)L.NY * k);
  double ue[24];
#pragma unroll
  for (int a = 0; a < 8; ++a) {
    const long long n = n0 + lnode_x(a) + (long long)L.NX * (lnode_y(a) + (long long)L.NY * lnode_z(a));
#pragma unroll
    for (int c = 0; c < 3; ++c) ue[3 * a + c] = nv[c * L.Np + n];
  }
  double s = 0.0;
#pragma unroll
  for (int r = 0; r < 24; ++r) {
    double t = 0.0;
#pragma unroll
    for (int c = 0; c < 24; ++c) t += c_KE0[r * 24 + c] * ue[c];
    s += ue[r] * t;
  }
  const double x = dens[e];
  const double xp = simp_pow(x, p);
  const double yo = (1.0 - xp) * y1 + xp * y2;
  const double dyo = p * pow(x, p - 1.0) * (y2 - y1);
  dfds[e] = -(dyo / yo) * (yo * s);
  dgds[e] = vol_e / vnorm;
}

// Element-mean stress and von Mises stress (reference problem/elasticity.cpp:
// 166-209).  The mean of B over the 2x2x2 Gauss points of a box voxel is B at
// its centre (B is multilinear), where dN_a/dX_j = s_aj / (4 h_j), s_aj = +-1
// the corner's side; so strain = B(0) u_e and stress = De(E(x_e)) strain.
__global__ __launch_bounds__(NT) void k_mises64(Lattice L, const double *__restrict__ dens,
                                                const double *__restrict__ nv, double y1,
                                                double y2, double p, double nu, double ihx,
                                                double ihy, double ihz,
                                                double *__restrict__ mises,
                                                double *__restrict__ stress) {
  const long long e = blockIdx.x * (long long)NT + threadIdx.x;
  if (e >= L.nelx) return;
  const int i = (int)(e % L.nx), j = (int)((e / L.nx) % L.ny), k = (int)(e / ((long long)L.nx * L.ny));
  const long long n0 = i + (long long)L.NX * (j + (long long)L.NY * k);
  // G[c][d] = du_c / dX_d at the centre
  double G[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
  const double ih[3] = {0.25 * ihx, 0.25 * ihy, 0.25 * ihz};
#pragma unroll
  for (int a = 0; a < 8; ++a) {
    const long long n = n0 + lnode_x(a) + (long long)L.NX * (lnode_y(a) + (long long)L.NY * lnode_z(a));
    const double sg[3] = {lnode_x(a) ? ih[0] : -ih[0], lnode_y(a) ? ih[1] : -ih[1],
                          lnode_z(a) ? ih[2] : -ih[2]};
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const double u = nv[c * L.Np + n];
#pragma unroll
      for (int d = 0; d < 3; ++d) G[c][d] += sg[d] * u;
    }
  }
  const double eps[6] = {G[0][0], G[1][1], G[2][2], G[0][1] + G[1][0], G[1][2] + G[2][1],
                         G[2][0] + G[0][2]};
  const double x = dens[e];
  const double xp = simp_pow(x, p);
  const double E = (1.0 - xp) * y1 + xp * y2;
  const double lam = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu));
  const double mu = E / (2.0 * (1.0 + nu));
  const double tr = eps[0] + eps[1] + eps[2];
  const double s[6] = {lam * tr + 2.0 * mu * eps[0], lam * tr + 2.0 * mu * eps[1],
                       lam * tr + 2.0 * mu * eps[2], mu * eps[3], mu * eps[4], mu * eps[5]};
  const double q = s[0] * s[0] + s[1] * s[1] + s[2] * s[2] - s[0] * s[1] - s[1] * s[2] -
                   s[2] * s[0] + 3.0 * (s[3] * s[3] + s[4] * s[4] + s[5] * s[5]);
  mises[e] = sqrt(fmax(q, 0.0));
  if (stress) {
#pragma unroll
    for (int c = 0; c < 6; ++c) stress[6 * e + c] = s[c];
  }
}

__global__ void k_numeq_to_lattice(Lattice L, const int *__restrict__ dof, int numeq,
                                   const double *__restrict__ v, const double *__restrict__ dirval,
                                   double *__restrict__ lat) {
  const long long n = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (n >= L.numnp) return;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const int q = dof[3 * n + c];
    lat[c * L.Np + n] = q < numeq ? v[q] : (dirval ? dirval[3 * n + c] : 0.0);
  }
}

__global__ void k_lattice_to_numeq(Lattice L, const int *__restrict__ dof, int numeq,
                                   const double *__restrict__ lat, double *__restrict__ v) {
  const long long n = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (n >= L.numnp) return;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const int q = dof[3 * n + c];
    if (q < numeq) v[q] = lat[c * L.Np + n];
  }
}

// y = x: the reference stream for HBM bandwidth and traffic counters
__global__ __launch_bounds__(NT) void k_copy64(long long n, const double *__restrict__ x,
                                               double *__restrict__ y) {
  const long long stride = (long long)gridDim.x * NT;
  for (long long i = blockIdx.x * (long long)NT + threadIdx.x; i < n; i += stride) y[i] = x[i];
}

__global__ void k_axpby(long long n, double a, const double *__restrict__ x, double b,
                        double *__restrict__ y) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i < n) y[i] = a * x[i] + b * y[i];
}

inline unsigned blocks_for(long long n, int nt) { return (unsigned)((n + nt - 1) / nt); }

}  // namespace

void launch_copy64(long long n, const double *x, double *y, hipStream_t s) {
  hipLaunchKernelGGL(k_copy64, dim3(VEC_BLOCKS), dim3(NT), 0, s, n, x, y);
  HIPCHK(hipGetLastError());
}

void launch_young(const double *x, double *young, long long nelx, double y1, double y2, double p,
                  hipStream_t s) {
  hipLaunchKernelGGL(k_young, dim3(blocks_for(nelx, 256)), dim3(256), 0, s, x, young, nelx, y1,
                     y2, p);
  HIPCHK(hipGetLastError());
}

void launch_diag_inv(const Lattice &L, const double *ke, const double *young,
                     const uint8_t *freemask, double *dinv, double *diag, hipStream_t s) {
  hipLaunchKernelGGL(k_diag_inv, dim3(blocks_for(L.numnp, 256)), dim3(256), 0, s, L, ke, young,
                     freemask, dinv, diag);
  HIPCHK(hipGetLastError());
}

static dim3 stencil_grid(const Lattice &L) {
  return dim3((L.NX + SX - 1) / SX, (L.NY + SY - 1) / SY, (L.NZ + SZ - 1) / SZ);
}

// Number of balanced z-chunks of the z-marching kernel: as many as make the
// columns x chunks one round of resident workgroups (2 per CU at the kernel's
// register use), at least 2 node planes per chunk.
static int resident_workgroups() {
  static int cus = 0;
  if (cus == 0) {
    int dev = 0;
    hipDeviceProp_t prop;
    cus = (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess &&
           prop.multiProcessorCount > 0)
              ? prop.multiProcessorCount
              : 256;
  }
  return 2 * cus;
}

int stencil_nchunks(const Lattice &L) {
  const int ncols = ((L.NX + ZX - 1) / ZX) * ((L.NY + ZY - 1) / ZY);
  return std::max(1, std::min(L.NZ / 2, resident_workgroups() / ncols));
}

// The split of a z-marching launch: one round of resident workgroups, each
// with at least 2 tile-planes; the autotuner's nchunks caps the blocks at
// tiles x nchunks.  The block count is a multiple of the tile rows where it
// can be, so every tile row is cut at the same planes: tiles of neighbouring
// rows (their 34-node y-halo planes are most of what tiles share) march in
// step through an XCD's L2.
static ZSplit make_zsplit(const Lattice &L, const KeCoef &ke) {
  ZSplit z;
  z.xlo = ke.xlo == 1 && L.NX > 1 ? 1 : 0;
  z.nbx = (L.NX - z.xlo + ZX - 1) / ZX;
  z.nby = (L.NY + ZY - 1) / ZY;
  const long long tiles = (long long)z.nbx * z.nby;
  z.W = tiles * L.NZ;
  long long nb = std::min<long long>(resident_workgroups(), z.W / 2);
  if (ke.nchunks > 0) nb = std::min<long long>(nb, tiles * ke.nchunks);
  if (nb >= z.nby) nb -= nb % z.nby;
  z.nb = (int)std::max<long long>(1, nb);
  return z;
}

long long stencil_blocks(const Lattice &L) {
  const dim3 a = stencil_grid(L);
  return std::max((long long)a.x * a.y * a.z, (long long)resident_workgroups());
}

KeCoef make_ke_coef_class(const double *ke, const double *ke_dev, int cls) {
  KeCoef k;
  k.ke = ke_dev;
  double kmax = 0.0;
  for (int e = 0; e < 576; ++e) kmax = std::max(kmax, std::fabs(ke[e]));
  auto fits = [&](const signed char *code, const int *rep, int n) {
    for (int e = 0; e < 576; ++e) {
      const int c = code[e];
      const double v = c > 0 ? ke[rep[c - 1]] : -ke[rep[-c - 1]];
      if (!(std::fabs(v - ke[e]) <= 1e-13 * kmax)) return false;
    }
    for (int i = 0; i < n; ++i) k.vals.v[i] = ke[rep[i]];
    return true;
  };
  if (cls == 2 && fits(KSYM_CUBE, KSYM_CUBE_REP, KSYM_CUBE_N)) k.sym = 2;
  else if (cls == 1 && fits(KSYM_BOX, KSYM_BOX_REP, KSYM_BOX_N)) k.sym = 1;
  for (int i = 0; i < 48; ++i) k.vals32.v[i] = (float)k.vals.v[i];
  return k;
}

KeCoef make_ke_coef(const double *ke, const double *ke_dev) {
  KeCoef k = make_ke_coef_class(ke, ke_dev, 2);
  if (k.sym == 0) k = make_ke_coef_class(ke, ke_dev, 1);
  return k;
}

template <bool PCG, int SYM, bool FUSE, int EPI = 0>
static void launch_stencil64_t(const Lattice &L, const KeCoef &ke, const double *young,
                               const uint8_t *freemask, const double *src, const double *dinv,
                               const double *p_old, double *p_new, double *y, PcgScalars *S,
                               double *partials, const double *pu, unsigned Gu, int j,
                               hipStream_t s, unsigned *nblocks, JacEpilogue je = JacEpilogue{}) {
  if (ke.kernel == 0) {
    const ZSplit z = make_zsplit(L, ke);
    hipLaunchKernelGGL((k_stencil64z<PCG, SYM, FUSE, EPI>), dim3((unsigned)z.nb), dim3(NT), 0, s, L,
                       z, ke.ke, ke.vals, young, freemask, src, dinv, p_old, p_new, y, S, partials,
                       pu, Gu, j, je);
    *nblocks = (unsigned)z.nb;
  } else {
    if (EPI) throw Error(kInternal, "the fused smoother needs the z-marching kernel");
    const dim3 g = stencil_grid(L);
    hipLaunchKernelGGL((k_stencil64<PCG, SYM, FUSE>), g, dim3(NT), 0, s, L, ke.ke, ke.vals, young,
                       freemask, src, dinv, p_old, p_new, y, S, partials, pu, Gu, j);
    *nblocks = g.x * g.y * g.z;
  }
  HIPCHK(hipGetLastError());
}

template <bool PCG, bool FUSE>
static unsigned launch_stencil64(const Lattice &L, const KeCoef &ke, const double *young,
                                 const uint8_t *freemask, const double *src, const double *dinv,
                                 const double *p_old, double *p_new, double *y, PcgScalars *S,
                                 double *partials, const double *pu, unsigned Gu, int j,
                                 hipStream_t s) {
  unsigned nb = 0;
  switch (ke.sym) {
    case 0:
      launch_stencil64_t<PCG, 0, FUSE>(L, ke, young, freemask, src, dinv, p_old, p_new, y, S,
                                       partials, pu, Gu, j, s, &nb);
      break;
    case 1:
      launch_stencil64_t<PCG, 1, FUSE>(L, ke, young, freemask, src, dinv, p_old, p_new, y, S,
                                       partials, pu, Gu, j, s, &nb);
      break;
    case 2:
      launch_stencil64_t<PCG, 2, FUSE>(L, ke, young, freemask, src, dinv, p_old, p_new, y, S,
                                       partials, pu, Gu, j, s, &nb);
      break;
    default:
      throw Error(kInvalidArgument, "unknown KE0 symmetry class");
  }
  return nb;
}

unsigned launch_apply_k64_jacobi(const Lattice &L, const KeCoef &ke, const double *young,
                                 const uint8_t *freemask, const double *x, double *xout,
                                 const double *b, const double *dinv, double omega,
                                 const double *q, double *partials, hipStream_t s) {
  if (ke.kernel != 0) throw Error(kInternal, "the fused smoother needs the z-marching kernel");
  if ((q == nullptr) != (partials == nullptr))
    throw Error(kInternal, "apply_k64_jacobi: q and partials go together");
  JacEpilogue je;
  je.b = b;
  je.q = q;
  je.omega = omega;
  unsigned nb = 0;
  // with the partials (the fine level's last sweep) or without (EPI 2)
#define JAC_LAUNCH(SYM_)                                                                          \
  if (partials)                                                                                   \
    launch_stencil64_t<false, SYM_, false, 1>(L, ke, young, freemask, x, dinv, nullptr, nullptr,  \
                                              xout, nullptr, partials, nullptr, 0, 0, s, &nb, je); \
  else                                                                                            \
    launch_stencil64_t<false, SYM_, false, 2>(L, ke, young, freemask, x, dinv, nullptr, nullptr,  \
                                              xout, nullptr, nullptr, nullptr, 0, 0, s, &nb, je);
  switch (ke.sym) {
    case 0: JAC_LAUNCH(0); break;
    case 1: JAC_LAUNCH(1); break;
    case 2: JAC_LAUNCH(2); break;
    default: throw Error(kInvalidArgument, "unknown KE0 symmetry class");
  }
#undef JAC_LAUNCH
  return nb;
}

void launch_apply_k64_wres(const Lattice &L, const KeCoef &ke, const double *young,
                           const uint8_t *freemask, const double *xc, const double *b,
                           const double *dinv, double omega, double *u, double *r, double *x0,
                           hipStream_t s) {
  if (ke.kernel != 0) throw Error(kInternal, "the fused W-cycle residual needs the z-marching K.u");
  if (!u || xc == u || xc == r || xc == x0) throw Error(kInternal, "apply_k64_wres: bad buffers");
  JacEpilogue je;
  je.b = b;
  je.omega = omega;
  je.u = u;
  je.x0 = x0;
  unsigned nb = 0;
  switch (ke.sym) {
    case 0:
      launch_stencil64_t<false, 0, false, 2>(L, ke, young, freemask, xc, dinv, nullptr, nullptr, r,
                                                nullptr, nullptr, nullptr, 0, 0, s, &nb, je);
      break;
    case 1:
      launch_stencil64_t<false, 1, false, 2>(L, ke, young, freemask, xc, dinv, nullptr, nullptr, r,
                                                nullptr, nullptr, nullptr, 0, 0, s, &nb, je);
      break;
    case 2:
      launch_stencil64_t<false, 2, false, 2>(L, ke, young, freemask, xc, dinv, nullptr, nullptr, r,
                                                nullptr, nullptr, nullptr, 0, 0, s, &nb, je);
      break;
    default:
      throw Error(kInvalidArgument, "unknown KE0 symmetry class");
  }
}

template <int EPI>
static unsigned launch_k32(const Lattice &L, const KeCoef &ke, const float *young32,
                           const uint8_t *freemask, const float *x, double *y,
                           const double *dinv, double *partials, JacEpilogue je, hipStream_t s) {
  if (ke.sym == 0 && !ke.ke32) throw Error(kInternal, "FP32 K.u without an FP32 KE0");
  const ZSplit z = make_zsplit(L, ke);
  const dim3 g((unsigned)z.nb);
  auto go = [&](auto kern) {
    hipLaunchKernelGGL(kern, g, dim3(NT), 0, s, L, z, ke.ke32, ke.vals32, young32, freemask, x,
                       dinv, nullptr, nullptr, y, nullptr, partials, nullptr, 0u, 0, je);
  };
  switch (ke.sym) {
    case 0: go(k_stencil64z<false, 0, false, EPI, float, double>); break;
    case 1: go(k_stencil64z<false, 1, false, EPI, float, double>); break;
    case 2: go(k_stencil64z<false, 2, false, EPI, float, double>); break;
    default: throw Error(kInvalidArgument, "unknown KE0 symmetry class");
  }
  HIPCHK(hipGetLastError());
  return g.x * g.y * g.z;
}

void launch_apply_k32(const Lattice &L, const KeCoef &ke, const float *young32,
                      const uint8_t *freemask, const float *x, double *y, hipStream_t s) {
  launch_k32<0>(L, ke, young32, freemask, x, y, nullptr, nullptr, JacEpilogue{}, s);
}

unsigned launch_apply_k32_jacobi(const Lattice &L, const KeCoef &ke, const float *young32,
                                 const uint8_t *freemask, const float *x, double *xout,
                                 const double *b, const double *dinv, double omega,
                                 const double *q, double *partials, hipStream_t s) {
  return launch_k32<1>(L, ke, young32, freemask, x, xout, dinv, partials,
                          JacEpilogue{b, q, omega}, s);
}

void launch_apply_k64(const Lattice &L, const KeCoef &ke, const double *young,
                      const uint8_t *freemask, const double *u, double *y, hipStream_t s) {
  launch_stencil64<false, false>(L, ke, young, freemask, u, nullptr, nullptr, nullptr, y, nullptr,
                                 nullptr, nullptr, 0, 0, s);
}

unsigned launch_pcg_stencil64(const Lattice &L, const KeCoef &ke, const double *young,
                              const uint8_t *freemask, const double *r, const double *dinv,
                              const double *p_old, double *p_new, double *q, PcgScalars *S,
                              double *partials, unsigned *counter, hipStream_t s, bool finalize) {
  (void)counter;
  const unsigned nb = launch_stencil64<true, false>(L, ke, young, freemask, r, dinv, p_old, p_new,
                                                    q, S, partials, nullptr, 0, 0, s);
  if (finalize) launch_finalize(kFinPcgAlpha, 1, partials, nb, S, nullptr, nullptr, s);
  return nb;
}

unsigned launch_pcg_stencil64_fused(const Lattice &L, const KeCoef &ke, const double *young,
                                    const uint8_t *freemask, const double *r, const double *dinv,
                                    const double *p_old, double *p_new, double *q, PcgScalars *S,
                                    double *partials_st, const double *partials_up, unsigned Gu,
                                    int j, hipStream_t s) {
  if (j == 0)
    return launch_stencil64<true, false>(L, ke, young, freemask, r, dinv, p_old, p_new, q, S,
                                         partials_st, nullptr, 0, 0, s);
  return launch_stencil64<true, true>(L, ke, young, freemask, r, dinv, p_old, p_new, q, S,
                                      partials_st, partials_up, Gu, j, s);
}

unsigned pcg_update_blocks() { return UPD_BLOCKS; }

unsigned launch_pcg_stencil64_mc(const Lattice &L, const KeCoef &ke, const double *young,
                                 const uint8_t *freemask, const double *r, const double *dinv,
                                 double *p, double *q, PcgScalars *S, double *partials_st,
                                 const double *partials_up, unsigned Gu, int j, int nc,
                                 long long cs, hipStream_t s) {
  if (ke.kernel != 0) throw Error(kInvalidArgument, "load cases together need the z-marching K.u");
  if (nc < 1) throw Error(kInvalidArgument, "no load cases");
  const ZSplit z = make_zsplit(L, ke);
  const unsigned nt = (unsigned)z.nb;
  const dim3 g(nt * (unsigned)nc);
#define MC_LAUNCH(SYM_)                                                                       \
  if (j == 0)                                                                                 \
    hipLaunchKernelGGL((k_stencil64z_mc<SYM_, false>), g, dim3(NT), 0, s, L, z, ke.ke,        \
                       ke.vals, young, freemask, r, dinv, p, q, S, partials_st, nullptr, 0u,  \
                       0, nc, cs);                                                            \
  else                                                                                        \
    hipLaunchKernelGGL((k_stencil64z_mc<SYM_, true>), g, dim3(NT), 0, s, L, z, ke.ke,         \
                       ke.vals, young, freemask, r, dinv, p, q, S, partials_st, partials_up,  \
                       Gu, j, nc, cs);
  switch (ke.sym) {
    case 0: MC_LAUNCH(0); break;
    case 1: MC_LAUNCH(1); break;
    case 2: MC_LAUNCH(2); break;
    default: throw Error(kInvalidArgument, "unknown KE0 symmetry class");
  }
#undef MC_LAUNCH
  HIPCHK(hipGetLastError());
  return nt;
}

unsigned launch_pcg_update64_mc(long long n, double *x, double *r, const double *p,
                                const double *q, const double *dinv, PcgScalars *S,
                                double *partials_up, const double *partials_st, unsigned Gs,
                                int j, int nc, long long cs, hipStream_t s) {
  hipLaunchKernelGGL(k_pcg_update64f_mc, dim3(UPD_BLOCKS, (unsigned)nc), dim3(NT), 0, s, n, x, r,
                     p, q, dinv, S, partials_up, partials_st, Gs, j, cs);
  HIPCHK(hipGetLastError());
  return UPD_BLOCKS;
}

unsigned launch_pcg_update64_fused(long long n, double *x, double *r, const double *p,
                                   const double *q, const double *dinv, PcgScalars *S,
                                   double *partials_up, const double *partials_st, unsigned Gs,
                                   int j, hipStream_t s) {
  hipLaunchKernelGGL(k_pcg_update64f, dim3(UPD_BLOCKS), dim3(NT), 0, s, n, x, r, p, q, dinv, S,
                     partials_up, partials_st, Gs, j);
  HIPCHK(hipGetLastError());
  return UPD_BLOCKS;
}

void launch_pcg_update64(long long n, double *x, double *r, const double *p, const double *q,
                         const double *dinv, PcgScalars *S, double *partials, unsigned *counter,
                         hipStream_t s, bool finalize) {
  hipLaunchKernelGGL(k_pcg_update64, dim3(VEC_BLOCKS), dim3(NT), 0, s, n, x, r, p, q, dinv, S,
                     partials, counter);
  HIPCHK(hipGetLastError());
  if (finalize) launch_finalize(kFinPcgBeta, 2, partials, VEC_BLOCKS, S, nullptr, nullptr, s);
}

void launch_pcg_init64(long long n, const double *b, double *x, double *r, double *p0,
                       const double *dinv, PcgScalars *S, double *partials, unsigned *counter,
                       hipStream_t s) {
  hipLaunchKernelGGL(k_pcg_init64, dim3(VEC_BLOCKS), dim3(NT), 0, s, n, b, x, r, p0, dinv, S,
                     partials, counter);
  HIPCHK(hipGetLastError());
  launch_finalize(kFinPcgInit, 2, partials, VEC_BLOCKS, S, nullptr, nullptr, s);
}

void launch_replace64(long long n, const double *b, const double *y, double *r,
                      const double *dinv, PcgScalars *S, double *partials, hipStream_t s) {
  hipLaunchKernelGGL(k_replace64, dim3(VEC_BLOCKS), dim3(NT), 0, s, n, b, y, r, dinv, partials);
  HIPCHK(hipGetLastError());
  launch_finalize(kFinReplace, 2, partials, VEC_BLOCKS, S, nullptr, nullptr, s);
}

void launch_residual_norms(long long n, const double *b, const double *y, double *out2,
                           double *partials, unsigned *counter, hipStream_t s) {
  hipLaunchKernelGGL(k_residual_norms, dim3(VEC_BLOCKS), dim3(NT), 0, s, n, b, y, out2, partials,
                     counter);
  HIPCHK(hipGetLastError());
  launch_finalize(kFinOut, 2, partials, VEC_BLOCKS, nullptr, nullptr, out2, s);
}

void launch_dot(long long n, const double *a, const double *b, double *out, double *partials,
                unsigned *counter, hipStream_t s) {
  hipLaunchKernelGGL(k_dot, dim3(VEC_BLOCKS), dim3(NT), 0, s, n, a, b, out, partials, counter);
  HIPCHK(hipGetLastError());
  launch_finalize(kFinOut, 1, partials, VEC_BLOCKS, nullptr, nullptr, out, s);
}

void launch_flush_small(long long n, const double *u, double *v, hipStream_t s) {
  hipLaunchKernelGGL(k_flush_small, dim3(blocks_for(n, 256)), dim3(256), 0, s, n, u, v);
  HIPCHK(hipGetLastError());
}

void launch_sensitivity64(const Lattice &L, const double *ke, const double *dens,
                          const double *nodeval, double y1, double y2, double p, double vol_e,
                          double vnorm, double *dfds, double *dgds, hipStream_t s) {
  hipLaunchKernelGGL(k_sensitivity64, dim3(blocks_for(L.nelx, NT)), dim3(NT), 0, s, L, ke, dens,
                     nodeval, y1, y2, p, vol_e, vnorm, dfds, dgds);
  HIPCHK(hipGetLastError());
}

void launch_mises64(const Lattice &L, const double *dens, const double *nodeval, double y1,
                    double y2, double p, double nu, const double h[3], double *mises,
                    double *stress, hipStream_t s) {
  hipLaunchKernelGGL(k_mises64, dim3(blocks_for(L.nelx, NT)), dim3(NT), 0, s, L, dens, nodeval, y1,
                     y2, p, nu, 1.0 / h[0], 1.0 / h[1], 1.0 / h[2], mises, stress);
  HIPCHK(hipGetLastError());
}

void launch_numeq_to_lattice(const Lattice &L, const int *dof, int numeq, const double *v,
                             const double *dirval, double *lat, hipStream_t s) {
  hipLaunchKernelGGL(k_numeq_to_lattice, dim3(blocks_for(L.numnp, 256)), dim3(256), 0, s, L, dof,
                     numeq, v, dirval, lat);
  HIPCHK(hipGetLastError());
}

void launch_lattice_to_numeq(const Lattice &L, const int *dof, int numeq, const double *lat,
                             double *v, hipStream_t s) {
  hipLaunchKernelGGL(k_lattice_to_numeq, dim3(blocks_for(L.numnp, 256)), dim3(256), 0, s, L, dof,
                     numeq, lat, v);
  HIPCHK(hipGetLastError());
}

void launch_axpby(long long n, double a, const double *x, double b, double *y, hipStream_t s) {
  hipLaunchKernelGGL(k_axpby, dim3(blocks_for(n, 256)), dim3(256), 0, s, n, a, x, b, y);
  HIPCHK(hipGetLastError());
}

// ---- load cases of the multigrid solve in one launch --------------------------
unsigned launch_pcg_stencil64_mcz(const Lattice &L, const KeCoef &ke, const double *young,
                                  const uint8_t *freemask, const double *z, const double *p_old,
                                  double *p_new, double *q, PcgScalars *S, double *partials,
                                  int nc, long long cs, hipStream_t s) {
  if (ke.kernel != 0) throw Error(kInvalidArgument, "load cases together need the z-marching K.u");
  if (nc < 1) throw Error(kInternal, "no load cases");
  const ZSplit z_ = make_zsplit(L, ke);
  const dim3 g((unsigned)z_.nb * (unsigned)nc);
  switch (ke.sym) {
    case 0: hipLaunchKernelGGL(k_stencil64z_mcz<0>, g, dim3(NT), 0, s, L, z_, ke.ke, ke.vals, young, freemask, z, p_old, p_new, q, S, partials, nc, cs); break;
    case 1: hipLaunchKernelGGL(k_stencil64z_mcz<1>, g, dim3(NT), 0, s, L, z_, ke.ke, ke.vals, young, freemask, z, p_old, p_new, q, S, partials, nc, cs); break;
    case 2: hipLaunchKernelGGL(k_stencil64z_mcz<2>, g, dim3(NT), 0, s, L, z_, ke.ke, ke.vals, young, freemask, z, p_old, p_new, q, S, partials, nc, cs); break;
    default: throw Error(kInvalidArgument, "unknown KE0 symmetry class");
  }
  HIPCHK(hipGetLastError());
  return (unsigned)z_.nb;
}

unsigned launch_apply_k64_mc(const Lattice &L, const KeCoef &ke, const double *young,
                             const uint8_t *freemask, const double *x, double *y, const double *b,
                             const double *dinv, double omega, const double *q, double *partials,
                             int nc, long long cs, hipStream_t s, double *u, double *x0) {
  if (ke.kernel != 0) throw Error(kInvalidArgument, "load cases together need the z-marching K.u");
  if (nc < 1) throw Error(kInternal, "no load cases");
  const int epi = b == nullptr ? 0 : partials ? 1 : 2;
  if (epi == 1 && (!q || !dinv)) throw Error(kInternal, "apply_k64_mc: the last sweep needs q");
  if (epi == 2 && !dinv) throw Error(kInternal, "apply_k64_mc: the sweep needs dinv");
  if ((u || x0) && (epi != 2 || !u || x == u || x == y || x == x0))
    throw Error(kInternal, "apply_k64_mc: bad W-cycle residual buffers");
  JacEpilogue je;
  je.b = b;
  je.q = epi == 1 ? q : nullptr;
  je.omega = omega;
  je.u = u;
  je.x0 = x0;
  const ZSplit z_ = make_zsplit(L, ke);
  const dim3 g((unsigned)z_.nb * (unsigned)nc);
#define MCV_LAUNCH(SYM_)                                                                           \
  if (epi == 1)                                                                                    \
    hipLaunchKernelGGL((k_stencil64z_mcv<SYM_, 1>), g, dim3(NT), 0, s, L, z_, ke.ke, ke.vals, young, \
                       freemask, x, dinv, y, partials, je, nc, cs);                               \
  else if (epi == 2)                                                                               \
    hipLaunchKernelGGL((k_stencil64z_mcv<SYM_, 2>), g, dim3(NT), 0, s, L, z_, ke.ke, ke.vals, young, \
                       freemask, x, dinv, y, nullptr, je, nc, cs);                                \
  else                                                                                             \
    hipLaunchKernelGGL((k_stencil64z_mcv<SYM_, 0>), g, dim3(NT), 0, s, L, z_, ke.ke, ke.vals, young, \
                       freemask, x, dinv, y, nullptr, je, nc, cs);
  switch (ke.sym) {
    case 0: MCV_LAUNCH(0); break;
    case 1: MCV_LAUNCH(1); break;
    case 2: MCV_LAUNCH(2); break;
    default: throw Error(kInvalidArgument, "unknown KE0 symmetry class");
  }
#undef MCV_LAUNCH
  HIPCHK(hipGetLastError());
  return (unsigned)z_.nb;
}

}  // namespace icarat_hip
