// k_learned.hip -- transfers between the node lattice and the U-Net's cell
// grid, and the flexible-PCG reductions of the learned preconditioner.
//
// Cell (I,J,K) of the network grid covers 2x2x2 elements; its centre is node
// (2I+1, 2J+1, 2K+1).  Restriction is full weighting over the 3x3x3 nodes
// around the centre (1-D weights 1/2, 1, 1/2; normalised by 8), prolongation
// is its transpose (trilinear interpolation from cell centres).
#include <hip/hip_runtime.h>

#include "learned_kernels.hpp"

namespace icarat_hip {

namespace {

constexpr int NT = 256;
constexpr int VEC_BLOCKS = 2048;

inline unsigned blocks_for(long long n, int nt) { return (unsigned)((n + nt - 1) / nt); }

__device__ __forceinline__ double w1(int d) { return d == 0 ? 1.0 : 0.5; }

// s[cell][c] = R(dinv r)  (scale 1); partial sums of s^2
__global__ __launch_bounds__(NT) void k_restrict_residual(Lattice L, NetGrid g, const double *__restrict__ r,
                                                          const double *__restrict__ dinv,
                                                          float *__restrict__ s3, double *partials) {
  __shared__ double sh[4];
  double acc = 0.0;
  const long long ncell = g.nvox();
  for (long long v = blockIdx.x * (long long)NT + threadIdx.x; v < ncell; v += (long long)gridDim.x * NT) {
    const int I = (int)(v % g.X), J = (int)((v / g.X) % g.Y), K = (int)(v / ((long long)g.X * g.Y));
    const int cx = 2 * I + 1, cy = 2 * J + 1, cz = 2 * K + 1;
    double s[3] = {0.0, 0.0, 0.0};
    for (int dz = -1; dz <= 1; ++dz) {
      const int z = cz + dz;
      if (z < 0 || z >= L.NZ) continue;
      for (int dy = -1; dy <= 1; ++dy) {
        const int y = cy + dy;
        if (y < 0 || y >= L.NY) continue;
        for (int dx = -1; dx <= 1; ++dx) {
          const int x = cx + dx;
          if (x < 0 || x >= L.NX) continue;
          const double w = w1(dx) * w1(dy) * w1(dz) * 0.125;
          const long long n = x + (long long)L.NX * (y + (long long)L.NY * z);
#pragma unroll
          for (int c = 0; c < 3; ++c) s[c] += w * dinv[c * L.Np + n] * r[c * L.Np + n];
        }
      }
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      s3[v * 3 + c] = (float)s[c];
      acc += s[c] * s[c];
    }
  }
  double vv[1] = {acc};
  store_partials<NT, 1>(vv, partials, sh);
}

// R(x) per cell (unscaled), for the training snapshots
__global__ void k_restrict_vec(Lattice L, NetGrid g, const double *__restrict__ x,
                               float *__restrict__ out3) {
  const long long v = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (v >= g.nvox()) return;
  const int I = (int)(v % g.X), J = (int)((v / g.X) % g.Y), K = (int)(v / ((long long)g.X * g.Y));
  const int cx = 2 * I + 1, cy = 2 * J + 1, cz = 2 * K + 1;
  double s[3] = {0.0, 0.0, 0.0};
  for (int dz = -1; dz <= 1; ++dz) {
    const int z = cz + dz;
    if (z < 0 || z >= L.NZ) continue;
    for (int dy = -1; dy <= 1; ++dy) {
      const int y = cy + dy;
      if (y < 0 || y >= L.NY) continue;
      for (int dx = -1; dx <= 1; ++dx) {
        const int xx = cx + dx;
        if (xx < 0 || xx >= L.NX) continue;
        const double w = w1(dx) * w1(dy) * w1(dz) * 0.125;
        const long long n = xx + (long long)L.NX * (y + (long long)L.NY * z);
#pragma unroll
        for (int c = 0; c < 3; ++c) s[c] += w * x[c * L.Np + n];
      }
    }
  }
#pragma unroll
  for (int c = 0; c < 3; ++c) out3[v * 3 + c] = (float)s[c];
}

// scale[0] = sigma = RMS of s over the cells, scale[1] = 1 / sigma (0 if sigma == 0)
__global__ void k_sigma(const double *sum2, long long ncell, double *scale) {
  const double sig = sqrt(sum2[0] / (double)(3 * ncell));
  scale[0] = sig;
  scale[1] = sig > 0.0 ? 1.0 / sig : 0.0;
}

// network input: [density feature, s / sigma (3), 0 x 4] in BF16
__global__ void k_pack_input(NetGrid g, const float *__restrict__ feat, const float *__restrict__ s3,
                             const double *__restrict__ scale, __bf16 *__restrict__ xin,
                             float *__restrict__ zr) {
  const long long v = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (v >= g.nvox()) return;
  const float inv = (float)scale[1];
  xin[v * 8 + 0] = (__bf16)feat[v];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const float a = s3[v * 3 + c] * inv;
    xin[v * 8 + 1 + c] = (__bf16)a;
    if (zr) zr[v * 3 + c] = a;
  }
#pragma unroll
  for (int c = 4; c < 8; ++c) xin[v * 8 + c] = (__bf16)0.0f;
}

// log-stiffness feature of each cell: (log10 mean young + 9) / 9 clamped to [0,1]
__global__ void k_density_feature(Lattice L, NetGrid g, const double *__restrict__ young,
                                  float *__restrict__ feat) {
  const long long v = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (v >= g.nvox()) return;
  const int I = (int)(v % g.X), J = (int)((v / g.X) % g.Y), K = (int)(v / ((long long)g.X * g.Y));
  double s = 0.0;
  int cnt = 0;
  for (int k = 0; k < 8; ++k) {
    const int ex = 2 * I + (k & 1), ey = 2 * J + ((k >> 1) & 1), ez = 2 * K + (k >> 2);
    if (ex < L.nx && ey < L.ny && ez < L.nz) {
      s += young[ex + (long long)L.nx * (ey + (long long)L.ny * ez)];
      ++cnt;
    }
  }
  float f = 0.f;
  if (cnt > 0 && s > 0.0) f = (float)((log10(s / cnt) + 9.0) / 9.0);
  feat[v] = fminf(fmaxf(f, 0.f), 1.f);
}

// z = dinv r + gain * sigma * P(out) on free DOFs
__global__ void k_prolong_add(Lattice L, NetGrid g, const float *__restrict__ out, int cs,
                              const double *__restrict__ scale, const double *__restrict__ r,
                              const double *__restrict__ dinv, const uint8_t *__restrict__ fmask,
                              double *__restrict__ z, int use_net, double gain,
                              double *__restrict__ corr_out) {
  const long long n = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (n >= L.numnp) return;
  const int x = (int)(n % L.NX), y = (int)((n / L.NX) % L.NY), zz = (int)(n / ((long long)L.NX * L.NY));
  double corr[3] = {0.0, 0.0, 0.0};
  if (use_net) {
    // cells whose centre is within one node of n, per dimension
    int ci[3][2], nc[3];
    double wc[3][2];
    const int co[3] = {x, y, zz};
    const int gd[3] = {g.X, g.Y, g.Z};
    for (int d = 0; d < 3; ++d) {
      nc[d] = 0;
      if (co[d] & 1) {
        const int c0 = (co[d] - 1) / 2;
        if (c0 < gd[d]) { ci[d][0] = c0; wc[d][0] = 1.0; nc[d] = 1; }
      } else {
        const int c0 = co[d] / 2 - 1, c1 = co[d] / 2;
        if (c0 >= 0 && c0 < gd[d]) { ci[d][nc[d]] = c0; wc[d][nc[d]] = 0.5; ++nc[d]; }
        if (c1 < gd[d]) { ci[d][nc[d]] = c1; wc[d][nc[d]] = 0.5; ++nc[d]; }
      }
    }
    for (int a = 0; a < nc[0]; ++a)
      for (int b = 0; b < nc[1]; ++b)
        for (int c = 0; c < nc[2]; ++c) {
          const double w = wc[0][a] * wc[1][b] * wc[2][c];
          const long long v = ci[0][a] + (long long)g.X * (ci[1][b] + (long long)g.Y * ci[2][c]);
#pragma unroll
          for (int k = 0; k < 3; ++k) corr[k] += w * (double)out[v * cs + k];
        }
  }
  const double sig = use_net ? gain * scale[0] : 0.0;
  const unsigned m = fmask[n];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const long long q = k * L.Np + n;
    const bool f = (m >> k) & 1u;
    z[q] = f ? dinv[q] * r[q] + sig * corr[k] : 0.0;
    if (corr_out) corr_out[q] = f ? sig * corr[k] : 0.0;
  }
}

// Descent guard of the corrected preconditioner z = b + c (b = dinv base):
// partials of b.r and c.r
__global__ __launch_bounds__(NT) void k_guard_dots(long long n, const double *__restrict__ base,
                                                   const double *__restrict__ dinv,
                                                   const double *__restrict__ c,
                                                   const double *__restrict__ r, double *partials) {
  __shared__ double sh[8];
  double br = 0.0, cr = 0.0;
  const long long stride = (long long)gridDim.x * NT;
  for (long long i = blockIdx.x * (long long)NT + threadIdx.x; i < n; i += stride) {
    br += dinv[i] * base[i] * r[i];
    cr += c[i] * r[i];
  }
  double v[2] = {br, cr};
  store_partials<NT, 2>(v, partials, sh);
}

// gamma in scale[3]: 1 when z.r = b.r + c.r >= b.r / 2, else the factor that
// makes it b.r / 2 (0 when b.r <= 0), so z.r stays positive
__global__ void k_guard_gamma(const double *__restrict__ partials, unsigned G, double *scale) {
  __shared__ double sh[8];
  double t[2];
  sum_partials<NT, 2>(partials, G, t, sh);
  if (threadIdx.x != 0) return;
  const double br = t[0], cr = t[1];
  double g = 1.0;
  if (!(br > 0.0)) g = 0.0;
  else if (br + cr < 0.5 * br) g = -0.5 * br / cr;
  scale[3] = g;
}

// z = b + gamma c where the guard shrank the correction
__global__ __launch_bounds__(NT) void k_guard_apply(long long n, const double *__restrict__ base,
                                                    const double *__restrict__ dinv,
                                                    const double *__restrict__ c,
                                                    const double *__restrict__ scale,
                                                    double *__restrict__ z) {
  const double g = scale[3];
  if (g == 1.0) return;
  const long long stride = (long long)gridDim.x * NT;
  for (long long i = blockIdx.x * (long long)NT + threadIdx.x; i < n; i += stride)
    z[i] = dinv[i] * base[i] + g * c[i];
}

__global__ void k_scale_f32(long long n, float a, float *__restrict__ x) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i < n) x[i] *= a;
}

// target = (R x* - R x_k) / sigma_k - zr_k
__global__ void k_make_target(long long n, const float *__restrict__ rxs, const float *__restrict__ rxk,
                              double inv_sigma, const float *__restrict__ zr, float *__restrict__ t) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i < n) t[i] = (float)(((double)rxs[i] - (double)rxk[i]) * inv_sigma) - zr[i];
}

// rho' = z.r and gamma = z.q (flexible PCG)
__global__ __launch_bounds__(NT) void k_zdot(long long n, const double *__restrict__ z,
                                             const double *__restrict__ r,
                                             const double *__restrict__ q, PcgScalars *S,
                                             double *partials) {
  __shared__ double sh[8];
  if (S->done) return;
  double a = 0.0, b = 0.0;
  const long long stride = (long long)gridDim.x * NT;
  for (long long i = blockIdx.x * (long long)NT + threadIdx.x; i < n; i += stride) {
    a += z[i] * r[i];
    b += z[i] * q[i];
  }
  double v[2] = {a, b};
  store_partials<NT, 2>(v, partials, sh);
}

__global__ void k_fpcg_beta(const double *__restrict__ partials, unsigned G, PcgScalars *S) {
  __shared__ double sh[8];
  if (S->done) return;
  double tot[2];
  sum_partials<NT, 2>(partials, G, tot, sh);
  if (threadIdx.x != 0) return;
  const double rho1 = tot[0], gam = tot[1];
  if (!(rho1 > 0.0) || !isfinite(rho1) || !isfinite(gam)) {
    S->status = 4;  // the preconditioned residual is not a descent direction
    S->done = 1;
    return;
  }
  // Polak-Ribiere: beta = z'.(r' - r) / rho with r - r' = alpha q
  S->beta = S->restart ? 0.0 : -S->alpha * gam / S->rho;
  S->restart = 0;
  S->rho = rho1;
}

// x += alpha p; r -= alpha q; r.r (FPCG: rho comes from the preconditioned residual)
__global__ __launch_bounds__(NT) void k_fpcg_update(long long n, double *__restrict__ x,
                                                    double *__restrict__ r,
                                                    const double *__restrict__ p,
                                                    const double *__restrict__ q, PcgScalars *S,
                                                    double *partials) {
  __shared__ double sh[8];
  if (S->done) return;
  const double alpha = S->alpha;
  double rr = 0.0;
  const long long stride = (long long)gridDim.x * NT;
  for (long long i = blockIdx.x * (long long)NT + threadIdx.x; i < n; i += stride) {
    x[i] += alpha * p[i];
    const double ri = r[i] - alpha * q[i];
    r[i] = ri;
    rr += ri * ri;
  }
  double v[1] = {rr};
  store_partials<NT, 1>(v, partials, sh);
}

// stop tests of the flexible PCG after the update (rr only)
__global__ void k_fpcg_check(const double *__restrict__ partials, unsigned G, PcgScalars *S) {
  __shared__ double sh[8];
  if (S->done) return;
  double tot[1];
  sum_partials<NT, 1>(partials, G, tot, sh);
  if (threadIdx.x != 0) return;
  S->rr = tot[0];
  S->iter += 1;
  const double rnorm = sqrt(tot[0]);
  if (!isfinite(rnorm)) {
    S->status = 2;
    S->done = 1;
  } else if (rnorm < S->tol * sqrt(S->bnorm2) || rnorm == 0.0) {
    S->status = 0;
    S->done = 1;
  } else if (S->iter > S->maxit) {
    S->status = 1;
    S->done = 1;
  } else if (rnorm < S->repl_norm) {
    S->status = 3;
    S->done = 1;
  }
}

// FPCG update with alpha folded in: every block sums the stencil's p.q
// partials in the same order, so all reach the same alpha = rho / p.q and the
// same breakdown decision; x += alpha p; r -= alpha q; r.r partials; with z0:
// z0 = omega dinv r as well (the first sweep of a V-cycle on r).
// COMP: x carries the low part xlo of a compensated sum (x + xlo is the exact
// sum of the updates; x is its rounding).  Near the FP64 floor of a
// high-contrast field the updates alpha p of the large displacements of
// near-void regions fall below half an ulp of x and would be lost, while the
// recursive residual counts them.
template <class TZ, bool COMP>
__global__ __launch_bounds__(NT) void k_fpcg_update_alpha(
    long long n, double *__restrict__ x, double *__restrict__ r, const double *__restrict__ p,
    const double *__restrict__ q, PcgScalars *S, const double *__restrict__ part_st, unsigned Gs,
    double *part_up, const double *__restrict__ dinv, double omega, TZ *__restrict__ z0,
    double *__restrict__ xlo, long long cs) {
  __shared__ double sh[8];
  // load case blockIdx.y: vectors at case x cs (dinv shared), its scalars,
  // the stencil's partials at case x Gs, ours at case x gridDim.x
  {
    const unsigned c = blockIdx.y;
    S += c;
    part_st += (size_t)c * Gs;
    part_up += (size_t)c * gridDim.x;
    x += c * cs;
    r += c * cs;
    p += c * cs;
    q += c * cs;
    if (z0) z0 += c * cs;
    if (COMP) xlo += c * cs;
  }
  if (S->done) return;
  double tot[1];
  sum_partials<NT, 1>(part_st, Gs, tot, sh);
  const double pq = tot[0];
  const double alpha = S->rho / pq;
  const bool ok = pq > 0.0 && isfinite(alpha);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    S->pq = pq;
    S->alpha = alpha;
    if (!ok) {
      S->status = 2;
      S->done = 1;
    }
  }
  if (!ok) return;
  double rr = 0.0;
  // one element: x (compensated with COMP), r, r.r, z0
  auto elem = [&](double &xi, double &li, double pi, double &ri, double qi, double di, TZ *zi) {
    if constexpr (COMP) {
      // two-sum of x and the update (with the carried low part)
      const double a = xi, b = fma(alpha, pi, li);
      const double sum = a + b, bb = sum - a;
      li = (a - (sum - bb)) + (b - bb);
      xi = sum;
    } else {
      xi += alpha * pi;
    }
    ri = ri - alpha * qi;
    rr += ri * ri;
    if (zi) *zi = (TZ)(omega * di * ri);
  };
  // pairs of elements with 16-byte accesses (the vectors are 16-byte aligned);
  // the streams no later kernel of the iteration reads (x and its low part,
  // p, q, dinv) bypass the caches' retention (non-temporal), r and z0 do not
  typedef double d2 __attribute__((ext_vector_type(2)));
  const long long half = n >> 1;
  const long long stride = (long long)gridDim.x * NT;
  for (long long i = blockIdx.x * (long long)NT + threadIdx.x; i < half; i += stride) {
    d2 xv = __builtin_nontemporal_load(reinterpret_cast<const d2 *>(x) + i);
    d2 rv = reinterpret_cast<const d2 *>(r)[i];
    const d2 pv = __builtin_nontemporal_load(reinterpret_cast<const d2 *>(p) + i);
    const d2 qv = __builtin_nontemporal_load(reinterpret_cast<const d2 *>(q) + i);
    d2 lv = {0.0, 0.0}, dv = {0.0, 0.0};
    if constexpr (COMP) lv = __builtin_nontemporal_load(reinterpret_cast<const d2 *>(xlo) + i);
    if (z0) dv = __builtin_nontemporal_load(reinterpret_cast<const d2 *>(dinv) + i);
    TZ zv[2];
    double x0 = xv.x, x1 = xv.y, l0 = lv.x, l1 = lv.y, r0 = rv.x, r1 = rv.y;
    elem(x0, l0, pv.x, r0, qv.x, dv.x, z0 ? &zv[0] : nullptr);
    elem(x1, l1, pv.y, r1, qv.y, dv.y, z0 ? &zv[1] : nullptr);
    __builtin_nontemporal_store(d2{x0, x1}, reinterpret_cast<d2 *>(x) + i);
    reinterpret_cast<d2 *>(r)[i] = d2{r0, r1};
    if constexpr (COMP) __builtin_nontemporal_store(d2{l0, l1}, reinterpret_cast<d2 *>(xlo) + i);
    if (z0) {
      z0[2 * i] = zv[0];
      z0[2 * i + 1] = zv[1];
    }
  }
  if ((n & 1) && blockIdx.x == 0 && threadIdx.x == 0) {
    const long long i = n - 1;
    double li = COMP ? xlo[i] : 0.0;
    TZ zi;
    elem(x[i], li, p[i], r[i], q[i], z0 ? dinv[i] : 0.0, z0 ? &zi : nullptr);
    if constexpr (COMP) xlo[i] = li;
    if (z0) z0[i] = zi;
  }
  double v[1] = {rr};
  store_partials_at<NT, 1>(v, part_up, sh, gridDim.x, blockIdx.x);
}

// z.r and z.q partials (no finalize); load case blockIdx.y: vectors at case x
// cs, its two partials per block at case x 2 gridDim.x
__global__ __launch_bounds__(NT) void k_zdot_partials(long long n, const double *__restrict__ z,
                                                      const double *__restrict__ r,
                                                      const double *__restrict__ q,
                                                      double *part_z, long long cs) {
  __shared__ double sh[8];
  {
    const unsigned c = blockIdx.y;
    z += c * cs;
    r += c * cs;
    q += c * cs;
    part_z += (size_t)c * 2 * gridDim.x;
  }
  double a = 0.0, b = 0.0;
  const long long stride = (long long)gridDim.x * NT;
  for (long long i = blockIdx.x * (long long)NT + threadIdx.x; i < n; i += stride) {
    a += z[i] * r[i];
    b += z[i] * q[i];
  }
  double v[2] = {a, b};
  store_partials_at<NT, 2>(v, part_z, sh, gridDim.x, blockIdx.x);
}

// One finalize per FPCG iteration: the stop tests of k_fpcg_check from the
// update's r.r partials, then (still running) rho' = z.r and the Polak-Ribiere
// beta of k_fpcg_beta from the z partials.
__global__ void k_fpcg_check_beta(const double *__restrict__ part_up, unsigned Gu,
                                  const double *__restrict__ part_z, unsigned Gz, PcgScalars *S) {
  __shared__ double sh[8];
  // one block per load case (one update partial, two z partials per block)
  part_up += (size_t)blockIdx.x * Gu;
  part_z += (size_t)blockIdx.x * 2 * Gz;
  S += blockIdx.x;
  if (S->done) return;
  double t1[1], t2[2];
  sum_partials<NT, 1>(part_up, Gu, t1, sh);
  sum_partials<NT, 2>(part_z, Gz, t2, sh);
  if (threadIdx.x != 0) return;
  S->rr = t1[0];
  S->iter += 1;
  const double rnorm = sqrt(t1[0]);
  if (!isfinite(rnorm)) {
    S->status = 2;
    S->done = 1;
    return;
  }
  if (rnorm < S->tol * sqrt(S->bnorm2) || rnorm == 0.0) {
    S->status = 0;
    S->done = 1;
    return;
  }
  if (S->iter > S->maxit) {
    S->status = 1;
    S->done = 1;
    return;
  }
  if (rnorm < S->repl_norm) {
    S->status = 3;
    S->done = 1;
    return;
  }
  const double rho1 = t2[0], gam = t2[1];
  if (!(rho1 > 0.0) || !isfinite(rho1) || !isfinite(gam)) {
    S->status = 4;
    S->done = 1;
    return;
  }
  S->beta = S->restart ? 0.0 : -S->alpha * gam / S->rho;
  S->restart = 0;
  S->rho = rho1;
}

// a restart of the flexible PCG (beta = 0 at the next beta step) unless done
__global__ void k_fpcg_restart_flag(PcgScalars *S) {
  if (threadIdx.x == 0 && !S->done) {
    S->restart = 1;
    S->alpha = 0.0;
  }
}

}  // namespace

void launch_restrict_residual(const Lattice &L, const NetGrid &g, const double *r, const double *dinv,
                              float *s3, double *partials, double *scale, hipStream_t s) {
  hipLaunchKernelGGL(k_restrict_residual, dim3(VEC_BLOCKS), dim3(NT), 0, s, L, g, r, dinv, s3,
                     partials);
  HIPCHK(hipGetLastError());
  launch_finalize(kFinOut, 1, partials, VEC_BLOCKS, nullptr, nullptr, scale + 2, s);
  hipLaunchKernelGGL(k_sigma, dim3(1), dim3(1), 0, s, scale + 2, g.nvox(), scale);
  HIPCHK(hipGetLastError());
}

void launch_restrict_vec(const Lattice &L, const NetGrid &g, const double *x, float *out3,
                         hipStream_t s) {
  hipLaunchKernelGGL(k_restrict_vec, dim3(blocks_for(g.nvox(), 256)), dim3(256), 0, s, L, g, x, out3);
  HIPCHK(hipGetLastError());
}

void launch_pack_input(const NetGrid &g, const float *feat, const float *s3, const double *scale,
                       __bf16 *xin, float *zr, hipStream_t s) {
  hipLaunchKernelGGL(k_pack_input, dim3(blocks_for(g.nvox(), 256)), dim3(256), 0, s, g, feat, s3,
                     scale, xin, zr);
  HIPCHK(hipGetLastError());
}

void launch_density_feature(const Lattice &L, const NetGrid &g, const double *young, float *feat,
                            hipStream_t s) {
  hipLaunchKernelGGL(k_density_feature, dim3(blocks_for(g.nvox(), 256)), dim3(256), 0, s, L, g,
                     young, feat);
  HIPCHK(hipGetLastError());
}

void launch_prolong_add(const Lattice &L, const NetGrid &g, const float *out, int cs,
                        const double *scale, const double *r, const double *dinv,
                        const uint8_t *fmask, double *z, int use_net, double gain, hipStream_t s,
                        double *corr_out) {
  hipLaunchKernelGGL(k_prolong_add, dim3(blocks_for(L.numnp, 256)), dim3(256), 0, s, L, g, out, cs,
                     scale, r, dinv, fmask, z, use_net, gain, corr_out);
  HIPCHK(hipGetLastError());
}

void launch_descent_guard(long long n, const double *base, const double *dinv, const double *c,
                          const double *r, double *z, double *partials, double *scale,
                          hipStream_t s) {
  hipLaunchKernelGGL(k_guard_dots, dim3(VEC_BLOCKS), dim3(NT), 0, s, n, base, dinv, c, r, partials);
  HIPCHK(hipGetLastError());
  hipLaunchKernelGGL(k_guard_gamma, dim3(1), dim3(NT), 0, s, partials, (unsigned)VEC_BLOCKS, scale);
  HIPCHK(hipGetLastError());
  hipLaunchKernelGGL(k_guard_apply, dim3(VEC_BLOCKS), dim3(NT), 0, s, n, base, dinv, c, scale, z);
  HIPCHK(hipGetLastError());
}

void launch_scale_f32(long long n, float a, float *x, hipStream_t s) {
  hipLaunchKernelGGL(k_scale_f32, dim3(blocks_for(n, 256)), dim3(256), 0, s, n, a, x);
  HIPCHK(hipGetLastError());
}

void launch_make_target(long long n, const float *rxs, const float *rxk, double inv_sigma,
                        const float *zr, float *t, hipStream_t s) {
  hipLaunchKernelGGL(k_make_target, dim3(blocks_for(n, 256)), dim3(256), 0, s, n, rxs, rxk,
                     inv_sigma, zr, t);
  HIPCHK(hipGetLastError());
}

void launch_zdot_beta(long long n, const double *z, const double *r, const double *q, PcgScalars *S,
                      double *partials, hipStream_t s) {
  hipLaunchKernelGGL(k_zdot, dim3(VEC_BLOCKS), dim3(NT), 0, s, n, z, r, q, S, partials);
  HIPCHK(hipGetLastError());
  hipLaunchKernelGGL(k_fpcg_beta, dim3(1), dim3(NT), 0, s, partials, (unsigned)VEC_BLOCKS, S);
  HIPCHK(hipGetLastError());
}

void launch_fpcg_update(long long n, double *x, double *r, const double *p, const double *q,
                        PcgScalars *S, double *partials, hipStream_t s) {
  hipLaunchKernelGGL(k_fpcg_update, dim3(VEC_BLOCKS), dim3(NT), 0, s, n, x, r, p, q, S, partials);
  HIPCHK(hipGetLastError());
  hipLaunchKernelGGL(k_fpcg_check, dim3(1), dim3(NT), 0, s, partials, (unsigned)VEC_BLOCKS, S);
  HIPCHK(hipGetLastError());
}

unsigned launch_fpcg_update_alpha(long long n, double *x, double *r, const double *p,
                                  const double *q, PcgScalars *S, const double *part_st,
                                  unsigned Gs, double *part_up, const double *dinv, double omega,
                                  double *z0, hipStream_t s, double *xlo, int nc, long long cs) {
  if (nc < 1) throw Error(kInternal, "fpcg update: no load case");
  const dim3 g(VEC_BLOCKS, (unsigned)nc);
  if (xlo)
    hipLaunchKernelGGL((k_fpcg_update_alpha<double, true>), g, dim3(NT), 0, s, n, x, r, p, q, S,
                       part_st, Gs, part_up, dinv, omega, z0, xlo, cs);
  else
    hipLaunchKernelGGL((k_fpcg_update_alpha<double, false>), g, dim3(NT), 0, s, n, x, r, p, q, S,
                       part_st, Gs, part_up, dinv, omega, z0, nullptr, cs);
  HIPCHK(hipGetLastError());
  return VEC_BLOCKS;
}

unsigned launch_fpcg_update_alpha(long long n, double *x, double *r, const double *p,
                                  const double *q, PcgScalars *S, const double *part_st,
                                  unsigned Gs, double *part_up, const double *dinv, double omega,
                                  float *z0, hipStream_t s, double *xlo) {
  if (xlo)
    hipLaunchKernelGGL((k_fpcg_update_alpha<float, true>), dim3(VEC_BLOCKS), dim3(NT), 0, s, n, x,
                       r, p, q, S, part_st, Gs, part_up, dinv, omega, z0, xlo, 0LL);
  else
    hipLaunchKernelGGL((k_fpcg_update_alpha<float, false>), dim3(VEC_BLOCKS), dim3(NT), 0, s, n, x,
                       r, p, q, S, part_st, Gs, part_up, dinv, omega, z0, nullptr, 0LL);
  HIPCHK(hipGetLastError());
  return VEC_BLOCKS;
}

unsigned launch_zdot_partials(long long n, const double *z, const double *r, const double *q,
                              double *part_z, hipStream_t s, int nc, long long cs) {
  if (nc < 1) throw Error(kInternal, "zdot partials: no load case");
  hipLaunchKernelGGL(k_zdot_partials, dim3(VEC_BLOCKS, (unsigned)nc), dim3(NT), 0, s, n, z, r, q,
                     part_z, cs);
  HIPCHK(hipGetLastError());
  return VEC_BLOCKS;
}

void launch_fpcg_restart_flag(PcgScalars *S, hipStream_t s) {
  hipLaunchKernelGGL(k_fpcg_restart_flag, dim3(1), dim3(64), 0, s, S);
  HIPCHK(hipGetLastError());
}

void launch_fpcg_check_beta(const double *part_up, unsigned Gu, const double *part_z, unsigned Gz,
                            PcgScalars *S, hipStream_t s, int nc) {
  if (nc < 1) throw Error(kInternal, "fpcg check: no load case");
  hipLaunchKernelGGL(k_fpcg_check_beta, dim3((unsigned)nc), dim3(NT), 0, s, part_up, Gu, part_z, Gz,
                     S);
  HIPCHK(hipGetLastError());
}

}  // namespace icarat_hip
