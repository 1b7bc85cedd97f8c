// k_unet.hip -- kernels of the learned preconditioner's 3-D U-Net: 3x3x3 /
// 1x1x1 convolutions as implicit GEMMs on BF16 MFMA (FP32 accumulation), the
// weight gradient on MFMA, pooling, up-sampling, channel concat/split, ReLU
// masks, the relative-L2 loss, Adam, and the restriction / prolongation
// between the node lattice and the network's cell grid.
//
// Activations are channels-last BF16: a[v*C + c], v = x + X*(y + Y*z).
// Weights are W[tap][co][ci] (BF16 copy of FP32 masters), tap = 9*(dz+1) +
// 3*(dy+1) + (dx+1).  The input gradient of a convolution is the same forward
// kernel run with the flipped, transposed weights Wt[26-tap][ci][co].
//
// MFMA used: v_mfma_f32_16x16x32_bf16; M = output channels, N = 16 voxels,
// K = 32 input channels of one tap (Cin is a multiple of 8, chunks of 32 are
// zero-padded at the end).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "unet_kernels.hpp"

namespace icarat_hip {

namespace {

typedef float v4f __attribute__((ext_vector_type(4)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));

constexpr int NT = 256;
constexpr int B = 4;                 // voxel block edge
constexpr int BV = B * B * B;        // 64 voxels per block
constexpr int HB = B + 2, HV = HB * HB * HB;  // halo block (216 voxels)

__device__ __forceinline__ int tap_dz(int t) { return t / 9 - 1; }
__device__ __forceinline__ int tap_dy(int t) { return (t / 3) % 3 - 1; }
__device__ __forceinline__ int tap_dx(int t) { return t % 3 - 1; }

struct Grid3 {
  int X, Y, Z;
  __device__ long long nvox() const { return (long long)X * Y * Z; }
};

__device__ __forceinline__ v8bf load8(const __bf16 *p) {
  return *reinterpret_cast<const v8bf *>(p);
}
__device__ __forceinline__ v8bf zero8() {
  v8bf z;
#pragma unroll
  for (int i = 0; i < 8; ++i) z[i] = (__bf16)0.0f;
  return z;
}

// Stages the (halo) block of x into LDS: rows of Cin bf16, zero outside.
template <int HALO>
__device__ __forceinline__ void stage_block(const Grid3 &g, int bx, int by, int bz, int Cin,
                                            const __bf16 *__restrict__ x, __bf16 *sx) {
  constexpr int E = B + 2 * HALO, EV = E * E * E;
  const int cv = Cin / 8;  // 16-byte chunks per voxel
  for (int t = threadIdx.x; t < EV * cv; t += NT) {
    const int row = t / cv, ch = t % cv;
    const int hx = row % E, hy = (row / E) % E, hz = row / (E * E);
    const int gx = bx * B - HALO + hx, gy = by * B - HALO + hy, gz = bz * B - HALO + hz;
    v8bf v = zero8();
    if (gx >= 0 && gx < g.X && gy >= 0 && gy < g.Y && gz >= 0 && gz < g.Z)
      v = load8(x + ((long long)gx + (long long)g.X * (gy + (long long)g.Y * gz)) * Cin + 8 * ch);
    *reinterpret_cast<v8bf *>(sx + row * Cin + 8 * ch) = v;
  }
}

// y = act(conv(x, W) + b).  One block of 4x4x4 voxels per workgroup, each
// wave 16 voxels x all output channels (MT row tiles of 16).
template <int TAPS, int MT>
__global__ __launch_bounds__(NT) void k_conv_fwd(Grid3 g, int Cin, int Cout,
                                                 const __bf16 *__restrict__ x,
                                                 const __bf16 *__restrict__ w,
                                                 const float *__restrict__ bias, int relu,
                                                 __bf16 *__restrict__ ybf, float *__restrict__ yf) {
  constexpr int HALO = TAPS == 27 ? 1 : 0;
  constexpr int E = B + 2 * HALO;
  extern __shared__ __align__(16) __bf16 sx[];
  const int nbx = (g.X + B - 1) / B, nby = (g.Y + B - 1) / B;
  const int bx = blockIdx.x % nbx, by = (blockIdx.x / nbx) % nby, bz = blockIdx.x / (nbx * nby);
  stage_block<HALO>(g, bx, by, bz, Cin, x, sx);
  __syncthreads();
  const int l = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int lv = wv * 16 + (l & 15);
  const int lx = lv % B, ly = (lv / B) % B, lz = lv / (B * B);
  v4f acc[MT];
#pragma unroll
  for (int m = 0; m < MT; ++m) acc[m] = v4f{0.f, 0.f, 0.f, 0.f};
  const int kq = 8 * (l >> 4);  // this lane's 8 channels within a 32-chunk
  for (int t = 0; t < TAPS; ++t) {
    const int dz = TAPS == 27 ? tap_dz(t) : 0, dy = TAPS == 27 ? tap_dy(t) : 0,
              dx = TAPS == 27 ? tap_dx(t) : 0;
    const int row = (lx + HALO + dx) + E * ((ly + HALO + dy) + E * (lz + HALO + dz));
    for (int c0 = 0; c0 < Cin; c0 += 32) {
      const int kc = c0 + kq;
      const v8bf bv = kc < Cin ? *reinterpret_cast<const v8bf *>(sx + row * Cin + kc) : zero8();
#pragma unroll
      for (int m = 0; m < MT; ++m) {
        const int co = 16 * m + (l & 15);
        const v8bf av = (kc < Cin && co < Cout)
                            ? load8(w + ((long long)t * Cout + co) * Cin + kc)
                            : zero8();
        acc[m] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv, acc[m], 0, 0, 0);
      }
    }
  }
  const int gx = bx * B + lx, gy = by * B + ly, gz = bz * B + lz;
  if (gx < g.X && gy < g.Y && gz < g.Z) {
    const long long v = gx + (long long)g.X * (gy + (long long)g.Y * gz);
#pragma unroll
    for (int m = 0; m < MT; ++m)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int co = 16 * m + 4 * (l >> 4) + r;
        if (co < Cout) {
          float val = acc[m][r] + (bias ? bias[co] : 0.f);
          if (relu) val = fmaxf(val, 0.f);
          if (ybf) ybf[v * Cout + co] = (__bf16)val;
          if (yf) yf[v * Cout + co] = val;
        }
      }
  }
}

// dW[tap][co][ci] += sum_v dA[v][co] x[v + off(tap)][ci] over this workgroup's
// share of the voxel blocks (atomic FP32 accumulation across workgroups).
// blockIdx.y = tap; tiles (co 16 x ci 16) are dealt to the 4 waves.
template <int TAPS>
__global__ __launch_bounds__(NT) void k_conv_wgrad(Grid3 g, int Cin, int Cout, int nparts,
                                                   const __bf16 *__restrict__ x,
                                                   const __bf16 *__restrict__ da,
                                                   float *__restrict__ dw) {
  constexpr int HALO = TAPS == 27 ? 1 : 0;
  constexpr int E = B + 2 * HALO;
  constexpr int TPW = 3;  // tiles per wave
  extern __shared__ __align__(16) __bf16 smem[];
  __bf16 *sx = smem;                              // E^3 x Cin
  __bf16 *sd = smem + E * E * E * Cin;            // 64 x Cout
  const int t = blockIdx.y;
  const int dz = TAPS == 27 ? tap_dz(t) : 0, dy = TAPS == 27 ? tap_dy(t) : 0,
            dx = TAPS == 27 ? tap_dx(t) : 0;
  const int mtc = (Cout + 15) / 16, ntc = (Cin + 15) / 16, ntiles = mtc * ntc;
  const int l = threadIdx.x & 63, wv = threadIdx.x >> 6;
  v4f acc[TPW];
#pragma unroll
  for (int k = 0; k < TPW; ++k) acc[k] = v4f{0.f, 0.f, 0.f, 0.f};
  const int nbx = (g.X + B - 1) / B, nby = (g.Y + B - 1) / B, nbz = (g.Z + B - 1) / B;
  const int nblocks = nbx * nby * nbz;
  for (int blk = blockIdx.x; blk < nblocks; blk += nparts) {
    const int bx = blk % nbx, by = (blk / nbx) % nby, bz = blk / (nbx * nby);
    __syncthreads();
    stage_block<HALO>(g, bx, by, bz, Cin, x, sx);
    for (int i = threadIdx.x; i < BV * Cout; i += NT) {
      const int lv = i / Cout, co = i % Cout;
      const int gx = bx * B + lv % B, gy = by * B + (lv / B) % B, gz = bz * B + lv / (B * B);
      __bf16 v = (__bf16)0.0f;
      if (gx < g.X && gy < g.Y && gz < g.Z)
        v = da[((long long)gx + (long long)g.X * (gy + (long long)g.Y * gz)) * Cout + co];
      sd[i] = v;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < TPW; ++k) {
      const int tile = wv * TPW + k;
      if (tile >= ntiles) break;
      const int mt = tile / ntc, nt = tile % ntc;
      const int co = 16 * mt + (l & 15), ci = 16 * nt + (l & 15);
#pragma unroll
      for (int ks = 0; ks < 2; ++ks) {
        v8bf av, bv;
#pragma unroll
        for (int r = 0; r < 8; ++r) {
          const int lv = 32 * ks + 8 * (l >> 4) + r;
          const int lx = lv % B, ly = (lv / B) % B, lz = lv / (B * B);
          const int row = (lx + HALO + dx) + E * ((ly + HALO + dy) + E * (lz + HALO + dz));
          av[r] = co < Cout ? sd[lv * Cout + co] : (__bf16)0.0f;
          bv[r] = ci < Cin ? sx[row * Cin + ci] : (__bf16)0.0f;
        }
        acc[k] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv, acc[k], 0, 0, 0);
      }
    }
  }
#pragma unroll
  for (int k = 0; k < TPW; ++k) {
    const int tile = wv * TPW + k;
    if (tile >= ntiles) break;
    const int mt = tile / ntc, nt = tile % ntc;
    const int ci = 16 * nt + (l & 15);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int co = 16 * mt + 4 * (l >> 4) + r;
      if (co < Cout && ci < Cin) atomicAdd(&dw[((long long)t * Cout + co) * Cin + ci], acc[k][r]);
    }
  }
}

__global__ void k_bias_grad(long long nvox, int C, const __bf16 *__restrict__ da,
                            float *__restrict__ db) {
  // one thread per (channel, voxel stripe)
  const int c = threadIdx.x % C;
  const int stripes = blockDim.x / C;
  const int s = threadIdx.x / C;
  if (s >= stripes) return;
  float acc = 0.f;
  for (long long v = blockIdx.x * (long long)stripes + s; v < nvox; v += (long long)gridDim.x * stripes)
    acc += (float)da[v * C + c];
  atomicAdd(&db[c], acc);
}

// da = dy * (y > 0)   (y = relu output)
__global__ void k_relu_mask(long long n, const __bf16 *__restrict__ y, const __bf16 *__restrict__ dy,
                            __bf16 *__restrict__ da) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i < n) da[i] = (float)y[i] > 0.f ? dy[i] : (__bf16)0.0f;
}

// 2x2x2 average pool (fine grid g -> g/2)
__global__ void k_pool(Grid3 gc, int C, const __bf16 *__restrict__ x, __bf16 *__restrict__ y) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  const long long n = gc.nvox() * C;
  if (i >= n) return;
  const int c = (int)(i % C);
  const long long v = i / C;
  const int X = gc.X, Y = gc.Y;
  const int cx = (int)(v % X), cy = (int)((v / X) % Y), cz = (int)(v / ((long long)X * Y));
  const long long FX = 2LL * X, FY = 2LL * Y;
  float s = 0.f;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const long long fv = (2 * cx + (k & 1)) + FX * ((2 * cy + ((k >> 1) & 1)) + FY * (2 * cz + (k >> 2)));
    s += (float)x[fv * C + c];
  }
  y[i] = (__bf16)(s * 0.125f);
}

// dx (+)= dy[parent] / 8
__global__ void k_pool_bwd(Grid3 gf, int C, const __bf16 *__restrict__ dy, __bf16 *__restrict__ dx,
                           int accumulate) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= gf.nvox() * C) return;
  const int c = (int)(i % C);
  const long long v = i / C;
  const int fx = (int)(v % gf.X), fy = (int)((v / gf.X) % gf.Y), fz = (int)(v / ((long long)gf.X * gf.Y));
  const long long pv = fx / 2 + (long long)(gf.X / 2) * (fy / 2 + (long long)(gf.Y / 2) * (fz / 2));
  float g = (float)dy[pv * C + c] * 0.125f;
  if (accumulate) g += (float)dx[i];
  dx[i] = (__bf16)g;
}

// nearest-neighbour x2 up-sampling into channels [off, off+C) of a Ctot tensor
__global__ void k_up(Grid3 gf, int C, int Ctot, int off, const __bf16 *__restrict__ x,
                     __bf16 *__restrict__ y) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= gf.nvox() * C) return;
  const int c = (int)(i % C);
  const long long v = i / C;
  const int fx = (int)(v % gf.X), fy = (int)((v / gf.X) % gf.Y), fz = (int)(v / ((long long)gf.X * gf.Y));
  const long long pv = fx / 2 + (long long)(gf.X / 2) * (fy / 2 + (long long)(gf.Y / 2) * (fz / 2));
  y[v * Ctot + off + c] = x[pv * C + c];
}

// dx[coarse][c] = sum of the 8 children dy[fine][off + c] (channels of a Ctot tensor)
__global__ void k_up_bwd(Grid3 gc, int C, int Ctot, int off, const __bf16 *__restrict__ dy,
                         __bf16 *__restrict__ dx) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= gc.nvox() * C) return;
  const int c = (int)(i % C);
  const long long v = i / C;
  const int X = gc.X, Y = gc.Y;
  const int cx = (int)(v % X), cy = (int)((v / X) % Y), cz = (int)(v / ((long long)X * Y));
  const long long FX = 2LL * X, FY = 2LL * Y;
  float s = 0.f;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const long long fv = (2 * cx + (k & 1)) + FX * ((2 * cy + ((k >> 1) & 1)) + FY * (2 * cz + (k >> 2)));
    s += (float)dy[fv * Ctot + off + c];
  }
  dx[i] = (__bf16)s;
}

// dst[v][doff + c] (+)= src[v][soff + c] for c < C
__global__ void k_copy_channels(long long nvox, int C, const __bf16 *__restrict__ src, int Cs,
                                int soff, __bf16 *__restrict__ dst, int Cd, int doff,
                                int accumulate) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= nvox * C) return;
  const int c = (int)(i % C);
  const long long v = i / C;
  float s = (float)src[v * Cs + soff + c];
  if (accumulate) s += (float)dst[v * Cd + doff + c];
  dst[v * Cd + doff + c] = (__bf16)s;
}

// relative L2 loss over the first cu of cs output channels against a
// cu-channel target: partial sums of (out - t)^2 and t^2; dout = (out - t) / sum t^2
__global__ void k_loss_terms(long long n, int cs, int cu, const float *__restrict__ out,
                             const float *__restrict__ t, double *acc2) {
  __shared__ double sh[8];
  double a = 0.0, b = 0.0;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const long long v = i / cu;
    const int c = (int)(i % cu);
    const double d = (double)out[v * cs + c] - (double)t[i];
    a += d * d;
    b += (double)t[i] * (double)t[i];
  }
  double v[2] = {a, b};
  block_sum<256, 2>(v, sh);
  if (threadIdx.x == 0) {
    atomicAdd(&acc2[0], v[0]);
    atomicAdd(&acc2[1], v[1]);
  }
}

__global__ void k_zero2(double *a) {
  a[0] = 0.0;
  a[1] = 0.0;
}

__global__ void k_loss_grad(long long n, int cs, int cu, const float *__restrict__ out,
                            const float *__restrict__ t, const double *acc2,
                            __bf16 *__restrict__ dout) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const long long v = i / cu;
  const int c = (int)(i % cu);
  const float inv = (float)(1.0 / acc2[1]);
  dout[v * cs + c] = (__bf16)((out[v * cs + c] - t[i]) * inv);
}

// Adam on FP32 masters; refreshes the BF16 weights and the flipped,
// transposed BF16 weights used for the input gradient.
__global__ void k_adam(long long n, int taps, int Cout, int Cin, float *__restrict__ p,
                       const float *__restrict__ g, float *__restrict__ m, float *__restrict__ v,
                       float lr, float b1, float b2, float eps, float c1, float c2,
                       __bf16 *__restrict__ wbf, __bf16 *__restrict__ wt) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float pi = p[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  p[i] = pi;
  if (wbf) {
    wbf[i] = (__bf16)pi;
    const int ci = (int)(i % Cin), co = (int)((i / Cin) % Cout), t = (int)(i / ((long long)Cin * Cout));
    wt[((long long)(taps - 1 - t) * Cin + ci) * Cout + co] = (__bf16)pi;
  }
}

__global__ void k_weights_bf16(long long n, int taps, int Cout, int Cin, const float *__restrict__ p,
                               __bf16 *__restrict__ wbf, __bf16 *__restrict__ wt) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float pi = p[i];
  wbf[i] = (__bf16)pi;
  const int ci = (int)(i % Cin), co = (int)((i / Cin) % Cout), t = (int)(i / ((long long)Cin * Cout));
  wt[((long long)(taps - 1 - t) * Cin + ci) * Cout + co] = (__bf16)pi;
}

__global__ void k_f32_to_bf16(long long n, const float *__restrict__ a, __bf16 *__restrict__ b) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i < n) b[i] = (__bf16)a[i];
}
__global__ void k_bf16_to_f32(long long n, const __bf16 *__restrict__ a, float *__restrict__ b) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i < n) b[i] = (float)a[i];
}

inline unsigned blocks_for(long long n, int nt) { return (unsigned)((n + nt - 1) / nt); }

}  // namespace

static Grid3 G3(const NetGrid &g) { return Grid3{g.X, g.Y, g.Z}; }

void launch_conv_fwd(const NetGrid &g, int taps, int Cin, int Cout, const __bf16 *x,
                     const __bf16 *w, const float *b, int relu, __bf16 *ybf, float *yf,
                     hipStream_t s) {
  if (Cin % 8 != 0 || Cin > 128) throw Error(kInvalidArgument, "conv: Cin must be a multiple of 8, <= 128");
  const unsigned nb = (unsigned)(((g.X + 3) / 4) * ((g.Y + 3) / 4) * ((g.Z + 3) / 4));
  const int E = taps == 27 ? 6 : 4;
  const size_t lds = (size_t)E * E * E * Cin * sizeof(__bf16);
  const int MT = (Cout + 15) / 16;
  Grid3 gg = G3(g);
#define CONV_CASE(T, M)                                                                          \
  hipLaunchKernelGGL((k_conv_fwd<T, M>), dim3(nb), dim3(NT), lds, s, gg, Cin, Cout, x, w, b, relu, \
                     ybf, yf)
  if (taps == 27) {
    switch (MT) {
      case 1: CONV_CASE(27, 1); break;
      case 2: CONV_CASE(27, 2); break;
      case 3: CONV_CASE(27, 3); break;
      case 4: CONV_CASE(27, 4); break;
      case 6: CONV_CASE(27, 6); break;
      default: throw Error(kInvalidArgument, "conv: unsupported Cout");
    }
  } else if (taps == 1) {
    switch (MT) {
      case 1: CONV_CASE(1, 1); break;
      case 2: CONV_CASE(1, 2); break;
      default: throw Error(kInvalidArgument, "conv1x1: unsupported Cout");
    }
  } else {
    throw Error(kInvalidArgument, "conv: taps must be 27 or 1");
  }
#undef CONV_CASE
  HIPCHK(hipGetLastError());
}

void launch_conv_wgrad(const NetGrid &g, int taps, int Cin, int Cout, const __bf16 *x,
                       const __bf16 *da, float *dw, hipStream_t s) {
  const int ntiles = ((Cout + 15) / 16) * ((Cin + 15) / 16);
  if (ntiles > 12) throw Error(kInvalidArgument, "wgrad: too many tiles");
  const int nblocks = ((g.X + 3) / 4) * ((g.Y + 3) / 4) * ((g.Z + 3) / 4);
  const int nparts = nblocks < 64 ? nblocks : 64;
  const int E = taps == 27 ? 6 : 4;
  const size_t lds = ((size_t)E * E * E * Cin + (size_t)BV * Cout) * sizeof(__bf16);
  Grid3 gg = G3(g);
  if (taps == 27)
    hipLaunchKernelGGL((k_conv_wgrad<27>), dim3(nparts, 27), dim3(NT), lds, s, gg, Cin, Cout,
                       nparts, x, da, dw);
  else
    hipLaunchKernelGGL((k_conv_wgrad<1>), dim3(nparts, 1), dim3(NT), lds, s, gg, Cin, Cout,
                       nparts, x, da, dw);
  HIPCHK(hipGetLastError());
}

void launch_bias_grad(const NetGrid &g, int C, const __bf16 *da, float *db, hipStream_t s) {
  const long long nv = (long long)g.X * g.Y * g.Z;
  const int per = 256 / C;
  hipLaunchKernelGGL(k_bias_grad, dim3(std::min<long long>(256, (nv + per - 1) / per)),
                     dim3(per * C), 0, s, nv, C, da, db);
  HIPCHK(hipGetLastError());
}

void launch_relu_mask(long long n, const __bf16 *y, const __bf16 *dy, __bf16 *da, hipStream_t s) {
  hipLaunchKernelGGL(k_relu_mask, dim3(blocks_for(n, 256)), dim3(256), 0, s, n, y, dy, da);
  HIPCHK(hipGetLastError());
}

void launch_pool(const NetGrid &gc, int C, const __bf16 *x, __bf16 *y, hipStream_t s) {
  const long long n = (long long)gc.X * gc.Y * gc.Z * C;
  hipLaunchKernelGGL(k_pool, dim3(blocks_for(n, 256)), dim3(256), 0, s, G3(gc), C, x, y);
  HIPCHK(hipGetLastError());
}

void launch_pool_bwd(const NetGrid &gf, int C, const __bf16 *dy, __bf16 *dx, int accumulate,
                     hipStream_t s) {
  const long long n = (long long)gf.X * gf.Y * gf.Z * C;
  hipLaunchKernelGGL(k_pool_bwd, dim3(blocks_for(n, 256)), dim3(256), 0, s, G3(gf), C, dy, dx,
                     accumulate);
  HIPCHK(hipGetLastError());
}

void launch_up(const NetGrid &gf, int C, int Ctot, int off, const __bf16 *x, __bf16 *y,
               hipStream_t s) {
  const long long n = (long long)gf.X * gf.Y * gf.Z * C;
  hipLaunchKernelGGL(k_up, dim3(blocks_for(n, 256)), dim3(256), 0, s, G3(gf), C, Ctot, off, x, y);
  HIPCHK(hipGetLastError());
}

void launch_up_bwd(const NetGrid &gc, int C, int Ctot, int off, const __bf16 *dy, __bf16 *dx,
                   hipStream_t s) {
  const long long n = (long long)gc.X * gc.Y * gc.Z * C;
  hipLaunchKernelGGL(k_up_bwd, dim3(blocks_for(n, 256)), dim3(256), 0, s, G3(gc), C, Ctot, off,
                     dy, dx);
  HIPCHK(hipGetLastError());
}

void launch_copy_channels(long long nvox, int C, const __bf16 *src, int Cs, int soff, __bf16 *dst,
                          int Cd, int doff, int accumulate, hipStream_t s) {
  const long long n = nvox * C;
  hipLaunchKernelGGL(k_copy_channels, dim3(blocks_for(n, 256)), dim3(256), 0, s, nvox, C, src, Cs,
                     soff, dst, Cd, doff, accumulate);
  HIPCHK(hipGetLastError());
}

void launch_loss(long long nvox, int cs, int cu, const float *out, const float *t, double *acc2,
                 __bf16 *dout, hipStream_t s) {
  const long long n = nvox * cu;
  hipLaunchKernelGGL(k_zero2, dim3(1), dim3(1), 0, s, acc2);
  hipLaunchKernelGGL(k_loss_terms, dim3(std::min<long long>(1024, blocks_for(n, 256))), dim3(256), 0,
                     s, n, cs, cu, out, t, acc2);
  hipLaunchKernelGGL(k_loss_grad, dim3(blocks_for(n, 256)), dim3(256), 0, s, n, cs, cu, out, t, acc2,
                     dout);
  HIPCHK(hipGetLastError());
}

void launch_adam(long long n, int taps, int Cout, int Cin, float *p, const float *g, float *m,
                 float *v, float lr, float b1, float b2, float eps, int step, __bf16 *wbf,
                 __bf16 *wt, hipStream_t s) {
  const float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
  hipLaunchKernelGGL(k_adam, dim3(blocks_for(n, 256)), dim3(256), 0, s, n, taps, Cout, Cin, p, g, m,
                     v, lr, b1, b2, eps, c1, c2, wbf, wt);
  HIPCHK(hipGetLastError());
}

void launch_weights_bf16(long long n, int taps, int Cout, int Cin, const float *p, __bf16 *wbf,
                         __bf16 *wt, hipStream_t s) {
  hipLaunchKernelGGL(k_weights_bf16, dim3(blocks_for(n, 256)), dim3(256), 0, s, n, taps, Cout, Cin,
                     p, wbf, wt);
  HIPCHK(hipGetLastError());
}

void launch_f32_to_bf16(long long n, const float *a, __bf16 *b, hipStream_t s) {
  hipLaunchKernelGGL(k_f32_to_bf16, dim3(blocks_for(n, 256)), dim3(256), 0, s, n, a, b);
  HIPCHK(hipGetLastError());
}

void launch_bf16_to_f32(long long n, const __bf16 *a, float *b, hipStream_t s) {
  hipLaunchKernelGGL(k_bf16_to_f32, dim3(blocks_for(n, 256)), dim3(256), 0, s, n, a, b);
  HIPCHK(hipGetLastError());
}

}  // namespace icarat_hip
