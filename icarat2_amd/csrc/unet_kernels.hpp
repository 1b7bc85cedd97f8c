// unet_kernels.hpp -- launchers of the learned preconditioner's network
// kernels (k_unet.hip) and of the lattice <-> cell-grid transfers (k_learned.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "device_common.hpp"
#include "hip_util.hpp"

namespace icarat_hip {

struct NetGrid {
  int X, Y, Z;
  __host__ __device__ long long nvox() const { return (long long)X * Y * Z; }
};

void launch_conv_fwd(const NetGrid &g, int taps, int Cin, int Cout, const __bf16 *x,
                     const __bf16 *w, const float *b, int relu, __bf16 *ybf, float *yf,
                     hipStream_t s);
void launch_conv_wgrad(const NetGrid &g, int taps, int Cin, int Cout, const __bf16 *x,
                       const __bf16 *da, float *dw, hipStream_t s);
void launch_bias_grad(const NetGrid &g, int C, const __bf16 *da, float *db, hipStream_t s);
void launch_relu_mask(long long n, const __bf16 *y, const __bf16 *dy, __bf16 *da, hipStream_t s);
void launch_pool(const NetGrid &gc, int C, const __bf16 *x, __bf16 *y, hipStream_t s);
void launch_pool_bwd(const NetGrid &gf, int C, const __bf16 *dy, __bf16 *dx, int accumulate,
                     hipStream_t s);
void launch_up(const NetGrid &gf, int C, int Ctot, int off, const __bf16 *x, __bf16 *y,
               hipStream_t s);
void launch_up_bwd(const NetGrid &gc, int C, int Ctot, int off, const __bf16 *dy, __bf16 *dx,
                   hipStream_t s);
void launch_copy_channels(long long nvox, int C, const __bf16 *src, int Cs, int soff, __bf16 *dst,
                          int Cd, int doff, int accumulate, hipStream_t s);
// relative L2 loss of the first cu of cs output channels against a cu-channel
// target; acc2 = {sum (out-t)^2, sum t^2}; dout (cs channels, the unused ones left untouched)
void launch_loss(long long nvox, int cs, int cu, const float *out, const float *t, double *acc2,
                 __bf16 *dout, hipStream_t s);
void launch_adam(long long n, int taps, int Cout, int Cin, float *p, const float *g, float *m,
                 float *v, float lr, float b1, float b2, float eps, int step, __bf16 *wbf,
                 __bf16 *wt, hipStream_t s);
void launch_weights_bf16(long long n, int taps, int Cout, int Cin, const float *p, __bf16 *wbf,
                         __bf16 *wt, hipStream_t s);
void launch_f32_to_bf16(long long n, const float *a, __bf16 *b, hipStream_t s);
void launch_bf16_to_f32(long long n, const __bf16 *a, float *b, hipStream_t s);

}  // namespace icarat_hip
