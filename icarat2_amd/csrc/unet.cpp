// unet.cpp -- host orchestration of the U-Net (see unet.hpp).
#include "unet.hpp"

#include <cmath>
#include <cstring>
#include <random>

namespace icarat_hip {

// c0..c5 (c5 padded to 8 output channels)
const UNetLayer UNet::kSpec[UNet::kLayers] = {
    {8, 16, 27}, {16, 32, 27}, {32, 64, 27}, {96, 32, 27}, {48, 16, 27}, {16, 8, 1}};

namespace {
template <class T>
void zero(DevBuf<T> &b, hipStream_t s) {
  if (b.n) HIPCHK(hipMemsetAsync(b.p, 0, b.bytes(), s));
}
int used_cout(int l) { return l == UNet::kLayers - 1 ? UNet::kOut : UNet::kSpec[l].cout; }
}  // namespace

UNet::UNet(const NetGrid &g0, hipStream_t s, uint64_t seed) : s_(s) {
  if (g0.X % 4 || g0.Y % 4 || g0.Z % 4) throw Error(kInvalidArgument, "U-Net grid must be a multiple of 4");
  g_[0] = g0;
  g_[1] = NetGrid{g0.X / 2, g0.Y / 2, g0.Z / 2};
  g_[2] = NetGrid{g0.X / 4, g0.Y / 4, g0.Z / 4};
  // He-normal initialisation from a project-owned generator
  std::mt19937_64 rng(seed);
  std::normal_distribution<float> nd(0.f, 1.f);
  for (int l = 0; l < kLayers; ++l) {
    const UNetLayer &L = kSpec[l];
    const long long nw = (long long)L.taps * L.cout * L.cin;
    Param &P = p_[l];
    for (DevBuf<float> *b : {&P.w, &P.gw, &P.mw, &P.vw}) { b->alloc(nw); zero(*b, s); }
    for (DevBuf<float> *b : {&P.b, &P.gb, &P.mb, &P.vb}) { b->alloc(L.cout); zero(*b, s); }
    P.wbf.alloc(nw);
    P.wt.alloc(nw);
    std::vector<float> w(nw, 0.f);
    const float std_ = std::sqrt(2.0f / (L.cin * L.taps));
    for (int t = 0; t < L.taps; ++t)
      for (int co = 0; co < used_cout(l); ++co)
        for (int ci = 0; ci < L.cin; ++ci) w[((long long)t * L.cout + co) * L.cin + ci] = nd(rng) * std_;
    HIPCHK(hipMemcpyAsync(P.w.p, w.data(), nw * sizeof(float), hipMemcpyHostToDevice, s));
    launch_weights_bf16(nw, L.taps, L.cout, L.cin, P.w, P.wbf, P.wt, s);
  }
  const long long v0 = g_[0].nvox(), v1 = g_[1].nvox(), v2 = g_[2].nvox();
  h0_.alloc(v0 * 16); p0_.alloc(v1 * 16); h1_.alloc(v1 * 32); p1_.alloc(v2 * 32);
  h2_.alloc(v2 * 64); k1_.alloc(v1 * 96); g1_.alloc(v1 * 32); k0_.alloc(v0 * 48);
  g0_.alloc(v0 * 16); out_.alloc(v0 * 8);
  dout_.alloc(v0 * 8); dg0_.alloc(v0 * 16); da_.alloc(v0 * 48); dk0_.alloc(v0 * 48);
  dg1_.alloc(v1 * 32); dh0_.alloc(v0 * 16); dk1_.alloc(v1 * 96); dh2_.alloc(v2 * 64);
  dh1_.alloc(v1 * 32); dp1_.alloc(v2 * 32); dp0_.alloc(v1 * 16);
  loss2_.alloc(2);
  zero(dout_, s);
  HIPCHK(hipStreamSynchronize(s));
}

long long UNet::host_param_count() {
  long long n = 0;
  for (int l = 0; l < kLayers; ++l) n += (long long)kSpec[l].taps * used_cout(l) * kSpec[l].cin + used_cout(l);
  return n;
}

void UNet::set_params(const float *host) {
  long long off = 0;
  for (int l = 0; l < kLayers; ++l) {
    const UNetLayer &L = kSpec[l];
    const int cu = used_cout(l);
    const long long nw = (long long)L.taps * L.cout * L.cin;
    std::vector<float> w(nw, 0.f), b(L.cout, 0.f);
    for (int t = 0; t < L.taps; ++t)
      for (int co = 0; co < cu; ++co)
        for (int ci = 0; ci < L.cin; ++ci)
          w[((long long)t * L.cout + co) * L.cin + ci] = host[off++];
    for (int co = 0; co < cu; ++co) b[co] = host[off++];
    HIPCHK(hipMemcpyAsync(p_[l].w.p, w.data(), nw * sizeof(float), hipMemcpyHostToDevice, s_));
    HIPCHK(hipMemcpyAsync(p_[l].b.p, b.data(), L.cout * sizeof(float), hipMemcpyHostToDevice, s_));
    launch_weights_bf16(nw, L.taps, L.cout, L.cin, p_[l].w, p_[l].wbf, p_[l].wt, s_);
    HIPCHK(hipStreamSynchronize(s_));
  }
}

// host layout: per layer the used weights (tap, co, ci) then the used biases
static void to_host(const DevBuf<float> &w, const DevBuf<float> &b, int l, float *host,
                    long long &off, hipStream_t s) {
  const UNetLayer &L = UNet::kSpec[l];
  const int cu = used_cout(l);
  const long long nw = (long long)L.taps * L.cout * L.cin;
  std::vector<float> hw(nw), hb(L.cout);
  HIPCHK(hipMemcpyAsync(hw.data(), w.p, nw * sizeof(float), hipMemcpyDeviceToHost, s));
  HIPCHK(hipMemcpyAsync(hb.data(), b.p, L.cout * sizeof(float), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  for (int t = 0; t < L.taps; ++t)
    for (int co = 0; co < cu; ++co)
      for (int ci = 0; ci < L.cin; ++ci) host[off++] = hw[((long long)t * L.cout + co) * L.cin + ci];
  for (int co = 0; co < cu; ++co) host[off++] = hb[co];
}

void UNet::get_params(float *host) const {
  long long off = 0;
  for (int l = 0; l < kLayers; ++l) to_host(p_[l].w, p_[l].b, l, host, off, s_);
}

void UNet::get_grads(float *host) const {
  long long off = 0;
  for (int l = 0; l < kLayers; ++l) to_host(p_[l].gw, p_[l].gb, l, host, off, s_);
}

long long UNet::flops_forward() const {
  long long f = 0;
  const long long v[3] = {g_[0].nvox(), g_[1].nvox(), g_[2].nvox()};
  const int lev[kLayers] = {0, 1, 2, 1, 0, 0};
  for (int l = 0; l < kLayers; ++l)
    f += 2LL * kSpec[l].taps * kSpec[l].cin * used_cout(l) * v[lev[l]];
  return f;
}

void UNet::forward(const __bf16 *x, float *out) {
  const NetGrid &G0 = g_[0], &G1 = g_[1], &G2 = g_[2];
  Param *P = p_;
  launch_conv_fwd(G0, 27, 8, 16, x, P[0].wbf, P[0].b, 1, h0_, nullptr, s_);
  launch_pool(G1, 16, h0_, p0_, s_);
  launch_conv_fwd(G1, 27, 16, 32, p0_, P[1].wbf, P[1].b, 1, h1_, nullptr, s_);
  launch_pool(G2, 32, h1_, p1_, s_);
  launch_conv_fwd(G2, 27, 32, 64, p1_, P[2].wbf, P[2].b, 1, h2_, nullptr, s_);
  launch_up(G1, 64, 96, 0, h2_, k1_, s_);
  launch_copy_channels(G1.nvox(), 32, h1_, 32, 0, k1_, 96, 64, 0, s_);
  launch_conv_fwd(G1, 27, 96, 32, k1_, P[3].wbf, P[3].b, 1, g1_, nullptr, s_);
  launch_up(G0, 32, 48, 0, g1_, k0_, s_);
  launch_copy_channels(G0.nvox(), 16, h0_, 16, 0, k0_, 48, 32, 0, s_);
  launch_conv_fwd(G0, 27, 48, 16, k0_, P[4].wbf, P[4].b, 1, g0_, nullptr, s_);
  launch_conv_fwd(G0, 1, 16, 8, g0_, P[5].wbf, P[5].b, 0, nullptr, out ? out : out_.p, s_);
  x_last_ = x;
}

void UNet::backward(const __bf16 *x) {
  const NetGrid &G0 = g_[0], &G1 = g_[1], &G2 = g_[2];
  Param *P = p_;
  for (int l = 0; l < kLayers; ++l) {
    zero(P[l].gw, s_);
    zero(P[l].gb, s_);
  }
  // c5 (1x1): dout (g0 x 8) -> dg0 (g0 x 16)
  launch_conv_wgrad(G0, 1, 16, 8, g0_, dout_, P[5].gw, s_);
  launch_bias_grad(G0, 8, dout_, P[5].gb, s_);
  launch_conv_fwd(G0, 1, 8, 16, dout_, P[5].wt, nullptr, 0, dg0_, nullptr, s_);
  // c4: relu mask, grads, dX -> dk0 (g0 x 48)
  launch_relu_mask(G0.nvox() * 16, g0_, dg0_, da_, s_);
  launch_conv_wgrad(G0, 27, 48, 16, k0_, da_, P[4].gw, s_);
  launch_bias_grad(G0, 16, da_, P[4].gb, s_);
  launch_conv_fwd(G0, 27, 16, 48, da_, P[4].wt, nullptr, 0, dk0_, nullptr, s_);
  // split k0 = [up(g1) | h0]
  launch_up_bwd(G1, 32, 48, 0, dk0_, dg1_, s_);
  launch_copy_channels(G0.nvox(), 16, dk0_, 48, 32, dh0_, 16, 0, 0, s_);
  // c3
  launch_relu_mask(G1.nvox() * 32, g1_, dg1_, da_, s_);
  launch_conv_wgrad(G1, 27, 96, 32, k1_, da_, P[3].gw, s_);
  launch_bias_grad(G1, 32, da_, P[3].gb, s_);
  launch_conv_fwd(G1, 27, 32, 96, da_, P[3].wt, nullptr, 0, dk1_, nullptr, s_);
  // split k1 = [up(h2) | h1]
  launch_up_bwd(G2, 64, 96, 0, dk1_, dh2_, s_);
  launch_copy_channels(G1.nvox(), 32, dk1_, 96, 64, dh1_, 32, 0, 0, s_);
  // c2
  launch_relu_mask(G2.nvox() * 64, h2_, dh2_, da_, s_);
  launch_conv_wgrad(G2, 27, 32, 64, p1_, da_, P[2].gw, s_);
  launch_bias_grad(G2, 64, da_, P[2].gb, s_);
  launch_conv_fwd(G2, 27, 64, 32, da_, P[2].wt, nullptr, 0, dp1_, nullptr, s_);
  launch_pool_bwd(G1, 32, dp1_, dh1_, 1, s_);
  // c1
  launch_relu_mask(G1.nvox() * 32, h1_, dh1_, da_, s_);
  launch_conv_wgrad(G1, 27, 16, 32, p0_, da_, P[1].gw, s_);
  launch_bias_grad(G1, 32, da_, P[1].gb, s_);
  launch_conv_fwd(G1, 27, 32, 16, da_, P[1].wt, nullptr, 0, dp0_, nullptr, s_);
  launch_pool_bwd(G0, 16, dp0_, dh0_, 1, s_);
  // c0
  launch_relu_mask(G0.nvox() * 16, h0_, dh0_, da_, s_);
  launch_conv_wgrad(G0, 27, 8, 16, x, da_, P[0].gw, s_);
  launch_bias_grad(G0, 16, da_, P[0].gb, s_);
}

void UNet::backward_only(const __bf16 *x, const float *target) {
  forward(x, nullptr);
  launch_loss(g_[0].nvox(), 8, kOut, out_, target, loss2_, dout_, s_);
  backward(x);
}

void UNet::eval_loss(const __bf16 *x, const float *target) {
  forward(x, nullptr);
  launch_loss(g_[0].nvox(), 8, kOut, out_, target, loss2_, dout_, s_);
}

void UNet::train_step(const __bf16 *x, const float *target, float lr) {
  backward_only(x, target);
  ++step_;
  for (int l = 0; l < kLayers; ++l) {
    const UNetLayer &L = kSpec[l];
    const long long nw = (long long)L.taps * L.cout * L.cin;
    launch_adam(nw, L.taps, L.cout, L.cin, p_[l].w, p_[l].gw, p_[l].mw, p_[l].vw, lr, 0.9f, 0.999f,
                1e-8f, step_, p_[l].wbf, p_[l].wt, s_);
    launch_adam(L.cout, 1, L.cout, 1, p_[l].b, p_[l].gb, p_[l].mb, p_[l].vb, lr, 0.9f, 0.999f,
                1e-8f, step_, nullptr, nullptr, s_);
  }
}

}  // namespace icarat_hip
