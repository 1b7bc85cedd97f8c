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

#include "kernels.hpp"

namespace icarat_hip {

namespace {

constexpr int SX = 32, SY = 8, SZ = 4, NT = 256;
constexpr int HX = SX + 2, HY = SY + 2, HZ = SZ + 2, HN = HX * HY * HZ;
constexpr int EX = SX + 1, EY = SY + 1, EZ = SZ + 1, EN = EX * EY * EZ;
constexpr int VEC_BLOCKS = 2048;

__device__ __forceinline__ constexpr int nb(int dx, int dy, int dz) {
  return (dx + 1) + 3 * (dy + 1) + 9 * (dz + 1);
}

// y = sum over the 8 elements around a node of E_s * KE0[rows of the node] . u_e
__device__ __forceinline__ void node_contract(const double *__restrict__ c_KE0,
                                              const double (&U)[27][3], const double (&E)[8],
                                              double (&y)[3]) {
  y[0] = y[1] = y[2] = 0.0;
#pragma unroll
  for (int s = 0; s < 8; ++s) {
    const int ex = (s & 1) - 1, ey = ((s >> 1) & 1) - 1, ez = ((s >> 2) & 1) - 1;
    const int a = lnode(-ex, -ey, -ez);
    double t0 = 0.0, t1 = 0.0, t2 = 0.0;
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const int o = nb(ex + lnode_x(b), ey + lnode_y(b), ez + lnode_z(b));
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        const double u = U[o][j];
        t0 += c_KE0[(3 * a + 0) * 24 + 3 * b + j] * u;
        t1 += c_KE0[(3 * a + 1) * 24 + 3 * b + j] * u;
        t2 += c_KE0[(3 * a + 2) * 24 + 3 * b + j] * u;
      }
    }
    y[0] += E[s] * t0;
    y[1] += E[s] * t1;
    y[2] += E[s] * t2;
  }
}

template <bool PCG>
__global__ __launch_bounds__(NT) void k_stencil64(
    Lattice L, const double *__restrict__ c_KE0, const double *__restrict__ young,
    const uint8_t *__restrict__ fmask, const double *__restrict__ src, const double *__restrict__ dinv,
    const double *__restrict__ p_old, double *__restrict__ p_new, double *__restrict__ y,
    PcgScalars *S, double *partials, unsigned *counter) {
  __shared__ double su[3][HN];
  __shared__ double se[EN];
  __shared__ double sh[4];
  double beta = 0.0;
  if constexpr (PCG) {
    if (S->done) return;
    beta = S->beta;
  }
  const int x0 = blockIdx.x * SX, y0 = blockIdx.y * SY, z0 = blockIdx.z * SZ;
  const long long Np = L.Np;
  for (int t = threadIdx.x; t < HN; t += NT) {
    const int hx = t % HX, hy = (t / HX) % HY, hz = t / (HX * HY);
    const int gx = x0 - 1 + hx, gy = y0 - 1 + hy, gz = z0 - 1 + hz;
    double v0 = 0.0, v1 = 0.0, v2 = 0.0;
    if (gx >= 0 && gx < L.NX && gy >= 0 && gy < L.NY && gz >= 0 && gz < L.NZ) {
      const long long n = gx + (long long)L.NX * (gy + (long long)L.NY * gz);
      if constexpr (PCG) {
        v0 = dinv[n] * src[n] + beta * p_old[n];
        v1 = dinv[Np + n] * src[Np + n] + beta * p_old[Np + n];
        v2 = dinv[2 * Np + n] * src[2 * Np + n] + beta * p_old[2 * Np + n];
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
      node_contract(c_KE0, U, E, out);
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
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double xp = pow(x[i], p);
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
  const double xp = pow(x, p);
  const double yo = (1.0 - xp) * y1 + xp * y2;
  const double dyo = p * pow(x, p - 1.0) * (y2 - y1);
  dfds[e] = -(dyo / yo) * (yo * s);
  dgds[e] = vol_e / vnorm;
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

__global__ void k_axpby(long long n, double a, const double *__restrict__ x, double b,
                        double *__restrict__ y) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i < n) y[i] = a * x[i] + b * y[i];
}

inline unsigned blocks_for(long long n, int nt) { return (unsigned)((n + nt - 1) / nt); }

}  // namespace

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

void launch_apply_k64(const Lattice &L, const double *ke, const double *young,
                      const uint8_t *freemask, const double *u, double *y, hipStream_t s) {
  hipLaunchKernelGGL(k_stencil64<false>, stencil_grid(L), dim3(NT), 0, s, L, ke, young, freemask, u,
                     nullptr, nullptr, nullptr, y, nullptr, nullptr, nullptr);
  HIPCHK(hipGetLastError());
}

void launch_pcg_stencil64(const Lattice &L, const double *ke, const double *young,
                          const uint8_t *freemask, const double *r, const double *dinv,
                          const double *p_old, double *p_new, double *q, PcgScalars *S,
                          double *partials, unsigned *counter, hipStream_t s, bool finalize) {
  const dim3 g = stencil_grid(L);
  hipLaunchKernelGGL(k_stencil64<true>, g, dim3(NT), 0, s, L, ke, young, freemask, r,
                     dinv, p_old, p_new, q, S, partials, counter);
  HIPCHK(hipGetLastError());
  if (finalize) launch_finalize(kFinPcgAlpha, 1, partials, g.x * g.y * g.z, S, nullptr, nullptr, s);
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

}  // namespace icarat_hip
