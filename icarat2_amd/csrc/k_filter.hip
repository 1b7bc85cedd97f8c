// k_filter.hip -- cone-weight density filter as an LDS-tiled 3-D stencil on
// the element lattice, the tanh projection, and the OC bisection step.
//
// The neighbour set of an element is the list of lattice offsets whose
// centre distance is <= rmin, in ascending element-id order ((dz,dy,dx)
// lexicographic), with weight rmin - distance (reference
// optimization/filter.hpp:254-297).  Sums run in that order with no FMA
// contraction, so results equal a sequential CPU evaluation bit for bit.
#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace icarat_hip {

namespace {

constexpr int FX = 32, FY = 8, FZ = 4, NT = 256;

template <int R>
struct Tile {
  static constexpr int HX = FX + 2 * R, HY = FY + 2 * R, HZ = FZ + 2 * R;
  static constexpr int HN = HX * HY * HZ;
};

// Loads a halo tile of `src` (zero outside the mesh); returns the inside mask
// in `msk`.  If F is given, the value stored is F(global element index).
template <int R, class F>
__device__ __forceinline__ void load_tile(const Lattice &L, int x0, int y0, int z0, double *tile,
                                          unsigned char *msk, F f) {
  using T = Tile<R>;
  for (int t = threadIdx.x; t < T::HN; t += NT) {
    const int hx = t % T::HX, hy = (t / T::HX) % T::HY, hz = t / (T::HX * T::HY);
    const int gx = x0 - R + hx, gy = y0 - R + hy, gz = z0 - R + hz;
    const bool in = gx >= 0 && gx < L.nx && gy >= 0 && gy < L.ny && gz >= 0 && gz < L.nz;
    tile[t] = in ? f(gx + (long long)L.nx * (gy + (long long)L.ny * gz)) : 0.0;
    if (msk) msk[t] = in ? 1 : 0;
  }
}

template <int R>
__global__ __launch_bounds__(NT) void k_filter(Lattice L, int count, const int *__restrict__ c_flin,
                                               const double *__restrict__ c_fw,
                                               const double *__restrict__ in,
                                               double *__restrict__ aa, double *__restrict__ out,
                                               int is_sens, double vol) {
#pragma clang fp contract(off)
  using T = Tile<R>;
  __shared__ double sv[T::HN];
  __shared__ double sa[T::HN];
  __shared__ unsigned char sm[T::HN];
  const int x0 = blockIdx.x * FX, y0 = blockIdx.y * FY, z0 = blockIdx.z * FZ;
  if (!is_sens) {
    load_tile<R>(L, x0, y0, z0, sv, sm, [&](long long e) { return in[e]; });
  } else {
    load_tile<R>(L, x0, y0, z0, sv, nullptr, [&](long long e) { return in[e]; });
    load_tile<R>(L, x0, y0, z0, sa, sm, [&](long long e) { return aa[e]; });
  }
  __syncthreads();
  const int tx = threadIdx.x % FX, ty = threadIdx.x / FX;
  for (int iz = 0; iz < FZ; ++iz) {
    const int gx = x0 + tx, gy = y0 + ty, gz = z0 + iz;
    if (gx >= L.nx || gy >= L.ny || gz >= L.nz) continue;
    const int base = (tx + R) + T::HX * ((ty + R) + T::HY * (iz + R));
    const long long e = gx + (long long)L.nx * (gy + (long long)L.ny * gz);
    if (!is_sens) {
      double a = 0.0, b = 0.0;
      for (int t = 0; t < count; ++t) {
        const int h = base + c_flin[t];
        const double w = c_fw[t];
        a += sm[h] ? w : 0.0;
        b += w * sv[h];
      }
      out[e] = b / a;
      aa[e] = a;
    } else if (is_sens == 1) {
      double s = 0.0;
      for (int t = 0; t < count; ++t) {
        const int h = base + c_flin[t];
        if (!sm[h]) continue;
        s += c_fw[t] * sv[h] / sa[h];
      }
      out[e] = s;
    } else {
      // sensitivity filter: sv = densities, sa = sensitivities
      double a = 0.0, b = 0.0;
      for (int t = 0; t < count; ++t) {
        const int h = base + c_flin[t];
        if (!sm[h]) continue;
        const double w = c_fw[t];
        a += w;
        b += (w * sv[h] * sa[h]) / (sv[h] * vol);
      }
      const double xi = sv[base];
      out[e] = ((xi * vol) / (xi * a)) * b;
    }
  }
}

__device__ __forceinline__ double oc_candidate(double x, double dc, double dv, double lam,
                                               double xmin, double xmax) {
  double r = -dc / (dv * lam);
  r = r > 0.0 ? r : 0.0;
  double v = sqrt(r) * x;
  const double lo = fmax(xmin, (1.0 - 0.15) * x);
  const double hi = fmin(xmax, (1.0 + 0.15) * x);
  if (v < lo) v = lo;
  else if (v > hi) v = hi;
  return v;
}

__device__ __forceinline__ double proj(double x, double beta, double T) {
  return 0.5 * (tanh(T * beta) + tanh(beta * (x - T))) / tanh(T * beta);
}

template <int R>
__global__ __launch_bounds__(NT) void k_oc_step(Lattice L, int count, const int *__restrict__ c_flin,
                                                const double *__restrict__ c_fw,
                                                const double *__restrict__ x,
                                                const double *__restrict__ dc,
                                                const double *__restrict__ dv, double xmin,
                                                double xmax, int use_projection, double beta,
                                                double vol_e, OcState *S, double *partials,
                                                unsigned *counter) {
#pragma clang fp contract(off)
  using T = Tile<R>;
  __shared__ double sv[T::HN];
  __shared__ unsigned char sm[T::HN];
  __shared__ double sh[4];
  if (S->done) return;
  const double lam = 0.5 * (S->l1 + S->l0);
  const int x0 = blockIdx.x * FX, y0 = blockIdx.y * FY, z0 = blockIdx.z * FZ;
  load_tile<R>(L, x0, y0, z0, sv, sm,
               [&](long long e) { return oc_candidate(x[e], dc[e], dv[e], lam, xmin, xmax); });
  __syncthreads();
  const int tx = threadIdx.x % FX, ty = threadIdx.x / FX;
  double vol = 0.0;
  for (int iz = 0; iz < FZ; ++iz) {
    const int gx = x0 + tx, gy = y0 + ty, gz = z0 + iz;
    if (gx >= L.nx || gy >= L.ny || gz >= L.nz) continue;
    const int base = (tx + R) + T::HX * ((ty + R) + T::HY * (iz + R));
    double a = 0.0, b = 0.0;
    for (int t = 0; t < count; ++t) {
      const int h = base + c_flin[t];
      const double w = c_fw[t];
      a += sm[h] ? w : 0.0;
      b += w * sv[h];
    }
    double xf = b / a;
    if (use_projection) xf = proj(xf, beta, 0.5);
    vol += xf * vol_e;
  }
  double v[1] = {vol};
  store_partials<NT, 1>(v, partials, sh);
}

__global__ void k_oc_final(long long n, double *__restrict__ x, const double *__restrict__ dc,
                           const double *__restrict__ dv, double xmin, double xmax,
                           const OcState *S) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  x[i] = oc_candidate(x[i], dc[i], dv[i], S->lam_last, xmin, xmax);
}

// x' = oc_candidate(x) at the bisection's current lambda (skipped once it is done)
__global__ void k_oc_candidate(long long n, const double *__restrict__ x,
                               const double *__restrict__ dc, const double *__restrict__ dv,
                               double xmin, double xmax, const OcState *S,
                               double *__restrict__ out) {
  if (S->done) return;
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  out[i] = oc_candidate(x[i], dc[i], dv[i], 0.5 * (S->l1 + S->l0), xmin, xmax);
}

// block partials of sum (projected) xf * vol_e for the kFinOc update
__global__ __launch_bounds__(NT) void k_oc_volume(long long n, const double *__restrict__ xf,
                                                  int use_projection, double beta, double vol_e,
                                                  const OcState *S, double *partials) {
#pragma clang fp contract(off)
  __shared__ double sh[4];
  if (S->done) return;
  double s = 0.0;
  const long long stride = (long long)gridDim.x * NT;
  for (long long i = blockIdx.x * (long long)NT + threadIdx.x; i < n; i += stride) {
    const double v = use_projection ? proj(xf[i], beta, 0.5) : xf[i];
    s += v * vol_e;
  }
  double v[1] = {s};
  store_partials<NT, 1>(v, partials, sh);
}

__global__ void k_threshold(long long n, const double *__restrict__ x, double beta, double T,
                            double *__restrict__ out, int deriv) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (!deriv) {
    out[i] = proj(x[i], beta, T);
  } else {
    const double t = tanh(beta * (x[i] - T));
    out[i] = (0.5 * (beta * (1.0 - t * t))) / tanh(T * beta);
  }
}

__global__ void k_mul(long long n, const double *__restrict__ a, double *__restrict__ b) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i < n) b[i] *= a[i];
}

__global__ __launch_bounds__(NT) void k_volume(long long n, const double *__restrict__ x,
                                               double vol_e, double *out, double *partials,
                                               unsigned *counter) {
  __shared__ double sh[4];
  double s = 0.0;
  const long long stride = (long long)gridDim.x * NT;
  for (long long i = blockIdx.x * (long long)NT + threadIdx.x; i < n; i += stride) s += x[i] * vol_e;
  double v[1] = {s};
  store_partials<NT, 1>(v, partials, sh);
}

inline unsigned blocks_for(long long n, int nt) { return (unsigned)((n + nt - 1) / nt); }

dim3 filter_grid(const Lattice &L) {
  return dim3((L.nx + FX - 1) / FX, (L.ny + FY - 1) / FY, (L.nz + FZ - 1) / FZ);
}

}  // namespace

void filter_lds_offsets(const int *off, int count, int R, int *lin) {
  const int HX = FX + 2 * R, HY = FY + 2 * R;
  for (int t = 0; t < count; ++t) lin[t] = off[3 * t] + HX * (off[3 * t + 1] + HY * off[3 * t + 2]);
}

#define FILTER_DISPATCH(R_, ...)                                             \
  switch (R_) {                                                              \
    case 1: { constexpr int RR = 1; __VA_ARGS__; } break;                    \
    case 2: { constexpr int RR = 2; __VA_ARGS__; } break;                    \
    case 3: { constexpr int RR = 3; __VA_ARGS__; } break;                    \
    default: throw Error(kInvalidArgument, "unsupported filter halo");       \
  }

void launch_density_filter(const Lattice &L, const FilterTable &T, const double *in, double *aa,
                           double *out, int is_sens, hipStream_t s) {
  FILTER_DISPATCH(T.R, hipLaunchKernelGGL(k_filter<RR>, filter_grid(L), dim3(NT), 0, s, L,
                                          T.count, T.lin, T.w, in, aa, out, is_sens, 0.0));
  HIPCHK(hipGetLastError());
}

void launch_sensitivity_filter(const Lattice &L, const FilterTable &T, const double *x,
                               const double *dfds, double *out, double vol_e, hipStream_t s) {
  FILTER_DISPATCH(T.R, hipLaunchKernelGGL(k_filter<RR>, filter_grid(L), dim3(NT), 0, s, L,
                                          T.count, T.lin, T.w, x, const_cast<double *>(dfds),
                                          out, 2, vol_e));
  HIPCHK(hipGetLastError());
}

void launch_oc_step(const Lattice &L, const FilterTable &T, const double *x, const double *dc,
                    const double *dv, double xmin, double xmax, int use_projection, double beta,
                    double vol_e, OcState *S, double *partials, unsigned *counter,
                    hipStream_t s) {
  FILTER_DISPATCH(T.R, hipLaunchKernelGGL(k_oc_step<RR>, filter_grid(L), dim3(NT), 0, s, L,
                                          T.count, T.lin, T.w, x, dc, dv, xmin, xmax,
                                          use_projection, beta,
                                          vol_e, S, partials, counter));
  HIPCHK(hipGetLastError());
  const dim3 g = filter_grid(L);
  launch_finalize(kFinOc, 1, partials, g.x * g.y * g.z, nullptr, S, nullptr, s);
}

void launch_oc_final(long long n, double *x, const double *dc, const double *dv, double xmin,
                     double xmax, const OcState *S, hipStream_t s) {
  hipLaunchKernelGGL(k_oc_final, dim3(blocks_for(n, 256)), dim3(256), 0, s, n, x, dc, dv, xmin,
                     xmax, S);
  HIPCHK(hipGetLastError());
}

void launch_oc_candidate(long long n, const double *x, const double *dc, const double *dv,
                         double xmin, double xmax, const OcState *S, double *out, hipStream_t s) {
  hipLaunchKernelGGL(k_oc_candidate, dim3(blocks_for(n, 256)), dim3(256), 0, s, n, x, dc, dv, xmin,
                     xmax, S, out);
  HIPCHK(hipGetLastError());
}

void launch_oc_volume(long long n, const double *xf, int use_projection, double beta, double vol_e,
                      OcState *S, double *partials, hipStream_t s) {
  constexpr unsigned G = 1024;
  hipLaunchKernelGGL(k_oc_volume, dim3(G), dim3(NT), 0, s, n, xf, use_projection, beta, vol_e, S,
                     partials);
  HIPCHK(hipGetLastError());
  launch_finalize(kFinOc, 1, partials, G, nullptr, S, nullptr, s);
}

void launch_threshold(long long n, const double *x, double beta, double T, double *out, int deriv,
                      hipStream_t s) {
  hipLaunchKernelGGL(k_threshold, dim3(blocks_for(n, 256)), dim3(256), 0, s, n, x, beta, T, out,
                     deriv);
  HIPCHK(hipGetLastError());
}

void launch_mul(long long n, const double *a, double *b, hipStream_t s) {
  hipLaunchKernelGGL(k_mul, dim3(blocks_for(n, 256)), dim3(256), 0, s, n, a, b);
  HIPCHK(hipGetLastError());
}

void launch_volume(long long n, const double *x, double vol_e, double *out, double *partials,
                   unsigned *counter, hipStream_t s) {
  hipLaunchKernelGGL(k_volume, dim3(2048), dim3(NT), 0, s, n, x, vol_e, out, partials, counter);
  HIPCHK(hipGetLastError());
  launch_finalize(kFinOut, 1, partials, 2048, nullptr, nullptr, out, s);
}

}  // namespace icarat_hip
