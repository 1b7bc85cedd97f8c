// unet.hpp -- the learned preconditioner's 3-D U-Net on the GPU: parameters
// (FP32 masters, Adam moments, BF16 working copies), activations, forward,
// backward and the Adam step, all on the project's own kernels.
#pragma once
#include <cstdint>
#include <vector>

#include "hip_util.hpp"
#include "unet_kernels.hpp"

namespace icarat_hip {

// Layer table (name order c0..c5): input channels, output channels, taps.
// The output layer has 8 channels on the GPU (3 used, 5 held at zero) so that
// every channel count is a multiple of 8.
struct UNetLayer {
  int cin, cout, taps;
};

class UNet {
 public:
  static constexpr int kLayers = 6;
  static constexpr int kIn = 8;    // input channels (4 used)
  static constexpr int kOut = 3;   // output channels used
  static const UNetLayer kSpec[kLayers];

  UNet(const NetGrid &g0, hipStream_t s, uint64_t seed);
  const NetGrid &grid(int level) const { return g_[level]; }
  // number of FP32 parameters in the host layout (w then b per layer, the
  // output layer with its 3 used channels)
  static long long host_param_count();
  void set_params(const float *host);
  void get_params(float *host) const;
  void get_grads(float *host) const;

  // out (g0 x 8 FP32, channels 0..2 used) = net(x) for x (g0 x 8 BF16)
  void forward(const __bf16 *x, float *out);
  // one training step on (x, target g0 x 3 FP32): forward, relative L2 loss,
  // backward, Adam.  The loss terms land in loss2[0..1] (device).
  void train_step(const __bf16 *x, const float *target, float lr);
  void backward_only(const __bf16 *x, const float *target);  // forward + loss + grads
  void eval_loss(const __bf16 *x, const float *target);      // forward + loss
  double *loss2() { return loss2_.p; }
  long long flops_forward() const;

 private:
  hipStream_t s_;
  NetGrid g_[3];
  struct Param {
    DevBuf<float> w, b, gw, gb, mw, vw, mb, vb;
    DevBuf<__bf16> wbf, wt;
  } p_[kLayers];
  int step_ = 0;
  DevBuf<__bf16> h0_, p0_, h1_, p1_, h2_, k1_, g1_, k0_, g0_;
  DevBuf<float> out_;
  DevBuf<__bf16> dout_, dg0_, da_, dk0_, dg1_, dh0_, dk1_, dh2_, dh1_, dp1_, dp0_, db_;
  DevBuf<double> loss2_;
  const __bf16 *x_last_ = nullptr;
  void backward(const __bf16 *x);
};

}  // namespace icarat_hip
