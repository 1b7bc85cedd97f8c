// k_galerkin.hip -- Galerkin coarse operators of the multigrid hierarchy:
// per-element 24x24 stiffness matrices K_C = sum_s W_s^T (M_f K_s M_f) W_s of
// every coarse voxel C from its 8 children s (the fine element matrices E_s KE0
// on the first coarse level, the explicit Galerkin matrices below it), masked
// by the free DOFs of both levels, so that the assembled coarse operator is
// exactly R A_f P with the multigrid's masked trilinear transfers.  W_s is the
// trilinear interpolation from C's 8 corners to the nodes of child s.  The
// coarse levels then smooth with these explicit matrices (one thread per
// node), and the coarsest dense matrix is assembled from them.
#include <hip/hip_runtime.h>

#include "galerkin_kernels.hpp"

namespace icarat_hip {

namespace {

constexpr int BT = 64;   // build kernel: one wave per coarse element
constexpr int NT = 256;

inline unsigned blocks_for(long long n, int nt) { return (unsigned)((n + nt - 1) / nt); }

// 1-D weight of coarse corner b (0/1) at fine offset f (0, 1, 2) inside C
__device__ __forceinline__ double w1(int f, int b) {
  return f == 1 ? 0.5 : ((f == 0) == (b == 0) ? 1.0 : 0.0);
}

// weight of coarse local node b at fine local node a of child s
__device__ __forceinline__ double wsab(int s, int a, int b) {
  const int fx = (s & 1) + lnode_x(a), fy = ((s >> 1) & 1) + lnode_y(a), fz = (s >> 2) + lnode_z(a);
  return w1(fx, lnode_x(b)) * w1(fy, lnode_y(b)) * w1(fz, lnode_z(b));
}

// FROM_FINE: children are E * KE0 (Kf = KE0, 576 entries; young per fine
// element); else children are explicit matrices Kf[e * 576 + ent].
template <bool FROM_FINE>
__global__ __launch_bounds__(BT) void k_gal_build(Lattice Lf, Lattice Lc,
                                                  const double *__restrict__ Kf,
                                                  const double *__restrict__ young,
                                                  const uint8_t *__restrict__ fmask_f,
                                                  const uint8_t *__restrict__ fmask_c,
                                                  double *__restrict__ Kc,
                                                  double *__restrict__ mu) {
  __shared__ double Ks[576], T[576], KC[576];
  __shared__ double sw[8][64];   // child s: w[a][b]
  __shared__ double dv[24], v[24], u[24];
  __shared__ unsigned char fm[24];
  const long long C = blockIdx.x;
  if (C >= Lc.nelx) return;
  const int cx = (int)(C % Lc.nx), cy = (int)((C / Lc.nx) % Lc.ny), cz = (int)(C / ((long long)Lc.nx * Lc.ny));
  const int t = threadIdx.x;
  for (int q = t; q < 512; q += BT) sw[q >> 6][q & 63] = wsab(q >> 6, (q >> 3) & 7, q & 7);
  for (int q = t; q < 576; q += BT) KC[q] = 0.0;
  __syncthreads();
  for (int s = 0; s < 8; ++s) {
    const int fx = 2 * cx + (s & 1), fy = 2 * cy + ((s >> 1) & 1), fz = 2 * cz + (s >> 2);
    const long long f = fx + (long long)Lf.nx * (fy + (long long)Lf.ny * fz);
    if (t < 24) {
      const int a = t / 3, c = t % 3;
      const long long n = (fx + lnode_x(a)) + (long long)Lf.NX * ((fy + lnode_y(a)) + (long long)Lf.NY * (fz + lnode_z(a)));
      fm[t] = (fmask_f[n] >> c) & 1u;
    }
    __syncthreads();
    const double E = FROM_FINE ? young[f] : 1.0;
    for (int q = t; q < 576; q += BT) {
      const int r = q / 24, cc = q % 24;
      const double k = FROM_FINE ? Kf[q] : Kf[f * 576 + q];
      Ks[q] = (fm[r] && fm[cc]) ? E * k : 0.0;
    }
    __syncthreads();
    // T = Ks W_s: T[(a c), (b c')] = sum_a' Ks[(a c), (a' c')] w[a'][b]
    for (int q = t; q < 576; q += BT) {
      const int r = q / 24, col = q % 24, b = col / 3, c2 = col % 3;
      double acc = 0.0;
#pragma unroll
      for (int a2 = 0; a2 < 8; ++a2) acc += Ks[r * 24 + 3 * a2 + c2] * sw[s][a2 * 8 + b];
      T[q] = acc;
    }
    __syncthreads();
    // KC += W_s^T T: KC[(b c), (b' c')] += sum_a w[a][b] T[(a c), (b' c')]
    for (int q = t; q < 576; q += BT) {
      const int r = q / 24, col = q % 24, b = r / 3, c = r % 3;
      double acc = 0.0;
#pragma unroll
      for (int a = 0; a < 8; ++a) acc += sw[s][a * 8 + b] * T[(3 * a + c) * 24 + col];
      KC[q] += acc;
    }
    __syncthreads();
  }
  // coarse masks, then the element matrix out (row-major, contiguous per element)
  if (t < 24) {
    const int a = t / 3, c = t % 3;
    const long long n = (cx + lnode_x(a)) + (long long)Lc.NX * ((cy + lnode_y(a)) + (long long)Lc.NY * (cz + lnode_z(a)));
    fm[t] = (fmask_c[n] >> c) & 1u;
  }
  __syncthreads();
  for (int q = t; q < 576; q += BT) {
    const int r = q / 24, cc = q % 24;
    if (!(fm[r] && fm[cc])) KC[q] = 0.0;
    Kc[C * 576 + q] = KC[q];
  }
  __syncthreads();
  // lambda_max(D^-1/2 K D^-1/2) over the element's free DOFs by power iteration
  if (t < 24) {
    const double d = KC[t * 24 + t];
    dv[t] = (fm[t] && d > 0.0) ? 1.0 / sqrt(d) : 0.0;
    v[t] = dv[t] != 0.0 ? 1.0 + 0.01 * t : 0.0;
  }
  __syncthreads();
  double lam = 0.0;
  for (int it = 0; it < 60; ++it) {
    if (t < 24) {
      double acc = 0.0;
#pragma unroll
      for (int j = 0; j < 24; ++j) acc += KC[t * 24 + j] * dv[j] * v[j];
      u[t] = dv[t] * acc;
    }
    __syncthreads();
    if (t == 0) {
      double nu = 0.0, nv = 0.0;
      for (int j = 0; j < 24; ++j) {
        nu += u[j] * u[j];
        nv += v[j] * v[j];
      }
      lam = nv > 0.0 ? sqrt(nu / nv) : 0.0;
      const double sc = nu > 0.0 ? 1.0 / sqrt(nu) : 0.0;
      for (int j = 0; j < 24; ++j) v[j] = u[j] * sc;
    }
    __syncthreads();
  }
  if (t == 0) mu[C] = lam;
}

// largest of n values into out[0] (one block)
__global__ __launch_bounds__(NT) void k_gal_max(long long n, const double *__restrict__ mu,
                                                double *__restrict__ out) {
  __shared__ double sh[NT / 64];
  double m = 0.0;
  for (long long i = threadIdx.x; i < n; i += NT) m = fmax(m, mu[i]);
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o, 64));
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double r = 0.0;
    for (int k = 0; k < NT / 64; ++k) r = fmax(r, sh[k]);
    out[0] = r;
  }
}

// dinv = theta / (mu_max diag) at free DOFs (the damped Jacobi factor), 0 elsewhere
__global__ __launch_bounds__(NT) void k_gal_diag(Lattice L, const double *__restrict__ Km,
                                                 const uint8_t *__restrict__ fmask,
                                                 const double *__restrict__ mu_max, double theta,
                                                 double *__restrict__ dinv) {
  const long long n = blockIdx.x * (long long)NT + threadIdx.x;
  if (n >= L.numnp) return;
  const int i = (int)(n % L.NX), j = (int)((n / L.NX) % L.NY), k = (int)(n / ((long long)L.NX * L.NY));
  double d[3] = {0.0, 0.0, 0.0};
  for (int s = 0; s < 8; ++s) {
    const int ex = i - 1 + (s & 1), ey = j - 1 + ((s >> 1) & 1), ez = k - 1 + (s >> 2);
    if (ex < 0 || ex >= L.nx || ey < 0 || ey >= L.ny || ez < 0 || ez >= L.nz) continue;
    const long long e = ex + (long long)L.nx * (ey + (long long)L.ny * ez);
    const int a = lnode(i - ex, j - ey, k - ez);
#pragma unroll
    for (int c = 0; c < 3; ++c) d[c] += Km[e * 576 + (3 * a + c) * 25];
  }
  const unsigned m = fmask[n];
  const double w = theta / mu_max[0];
#pragma unroll
  for (int c = 0; c < 3; ++c) dinv[c * L.Np + n] = (((m >> c) & 1u) && d[c] > 0.0) ? w / d[c] : 0.0;
}

// y = K x with the explicit element matrices, one thread per node; MODE:
// GAL_APPLY out = K x, GAL_RESIDUAL out = b - K x, GAL_JACOBI out = x + dinv (b - K x)
// (out may not alias x); every output masked by the free DOFs.
template <int MODE>
__global__ __launch_bounds__(NT) void k_gal_apply(Lattice L, const double *__restrict__ Km,
                                                  const uint8_t *__restrict__ fmask,
                                                  const double *__restrict__ x,
                                                  const double *__restrict__ b,
                                                  const double *__restrict__ dinv,
                                                  double *__restrict__ out) {
  const long long n = blockIdx.x * (long long)NT + threadIdx.x;
  if (n >= L.numnp) return;
  const int i = (int)(n % L.NX), j = (int)((n / L.NX) % L.NY), k = (int)(n / ((long long)L.NX * L.NY));
  double y[3] = {0.0, 0.0, 0.0};
  for (int s = 0; s < 8; ++s) {
    const int ex = i - 1 + (s & 1), ey = j - 1 + ((s >> 1) & 1), ez = k - 1 + (s >> 2);
    if (ex < 0 || ex >= L.nx || ey < 0 || ey >= L.ny || ez < 0 || ez >= L.nz) continue;
    const long long e = ex + (long long)L.nx * (ey + (long long)L.ny * ez);
    const int a = lnode(i - ex, j - ey, k - ez);
    const double *__restrict__ row = Km + e * 576 + 3 * a * 24;
#pragma unroll
    for (int bb = 0; bb < 8; ++bb) {
      const long long nb = (ex + lnode_x(bb)) + (long long)L.NX * ((ey + lnode_y(bb)) + (long long)L.NY * (ez + lnode_z(bb)));
#pragma unroll
      for (int c2 = 0; c2 < 3; ++c2) {
        const double u = x[c2 * L.Np + nb];
#pragma unroll
        for (int c = 0; c < 3; ++c) y[c] += row[c * 24 + 3 * bb + c2] * u;
      }
    }
  }
  const unsigned m = fmask[n];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const long long q = c * L.Np + n;
    const bool fr = (m >> c) & 1u;
    double o;
    if constexpr (MODE == GAL_APPLY) o = fr ? y[c] : 0.0;
    else if constexpr (MODE == GAL_RESIDUAL) o = fr ? b[q] - y[c] : 0.0;
    else o = fr ? x[q] + dinv[q] * (b[q] - y[c]) : x[q];
    out[q] = o;
  }
}

// dense assembly of one parity class of elements from the explicit matrices
__global__ void k_gal_dense_scatter(Lattice L, const double *__restrict__ Km, int px, int py,
                                    int pz, double *__restrict__ A) {
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
  const long long e = ex + (long long)L.nx * (ey + (long long)L.ny * ez);
  A[(i * L.numnp + na) * n + j * L.numnp + nb_] += Km[e * 576 + ent];
}

}  // namespace

void launch_gal_build(const Lattice &Lf, const Lattice &Lc, const double *kf, const double *young,
                      const uint8_t *fmask_f, const uint8_t *fmask_c, double *kc, double *mu,
                      hipStream_t s) {
  if (Lf.nx != 2 * Lc.nx || Lf.ny != 2 * Lc.ny || Lf.nz != 2 * Lc.nz)
    throw Error(kInternal, "Galerkin build needs a 2x coarsening");
  if (young)
    hipLaunchKernelGGL(k_gal_build<true>, dim3((unsigned)Lc.nelx), dim3(BT), 0, s, Lf, Lc, kf,
                       young, fmask_f, fmask_c, kc, mu);
  else
    hipLaunchKernelGGL(k_gal_build<false>, dim3((unsigned)Lc.nelx), dim3(BT), 0, s, Lf, Lc, kf,
                       nullptr, fmask_f, fmask_c, kc, mu);
  HIPCHK(hipGetLastError());
}

void launch_gal_dinv(const Lattice &L, const double *km, const uint8_t *fmask, const double *mu,
                     double *mu_max, double theta, double *dinv, hipStream_t s) {
  hipLaunchKernelGGL(k_gal_max, dim3(1), dim3(NT), 0, s, L.nelx, mu, mu_max);
  hipLaunchKernelGGL(k_gal_diag, dim3(blocks_for(L.numnp, NT)), dim3(NT), 0, s, L, km, fmask,
                     mu_max, theta, dinv);
  HIPCHK(hipGetLastError());
}

void launch_gal_apply(int mode, const Lattice &L, const double *km, const uint8_t *fmask,
                      const double *x, const double *b, const double *dinv, double *out,
                      hipStream_t s) {
  const dim3 g(blocks_for(L.numnp, NT)), t(NT);
  switch (mode) {
    case GAL_APPLY:
      hipLaunchKernelGGL(k_gal_apply<GAL_APPLY>, g, t, 0, s, L, km, fmask, x, b, dinv, out);
      break;
    case GAL_RESIDUAL:
      hipLaunchKernelGGL(k_gal_apply<GAL_RESIDUAL>, g, t, 0, s, L, km, fmask, x, b, dinv, out);
      break;
    case GAL_JACOBI:
      hipLaunchKernelGGL(k_gal_apply<GAL_JACOBI>, g, t, 0, s, L, km, fmask, x, b, dinv, out);
      break;
    default:
      throw Error(kInvalidArgument, "unknown Galerkin apply mode");
  }
  HIPCHK(hipGetLastError());
}

void launch_gal_dense_scatter(const Lattice &L, const double *km, double *A, hipStream_t s) {
  for (int cls = 0; cls < 8; ++cls) {
    const int px = cls & 1, py = (cls >> 1) & 1, pz = cls >> 2;
    const long long cnt = (long long)((L.nx - px + 1) / 2) * ((L.ny - py + 1) / 2) * ((L.nz - pz + 1) / 2) * 576;
    if (cnt == 0) continue;
    hipLaunchKernelGGL(k_gal_dense_scatter, dim3(blocks_for(cnt, 256)), dim3(256), 0, s, L, km, px,
                       py, pz, A);
  }
  HIPCHK(hipGetLastError());
}

}  // namespace icarat_hip
