// learned.cpp -- the learned preconditioner (see learned.hpp).
#include "learned.hpp"

#include <algorithm>
#include <cmath>

#include "context.hpp"
#include "learned_kernels.hpp"
#include "mixed_kernels.hpp"
#include "multigrid.hpp"

namespace icarat_hip {

namespace {
int round4(int v) { return (v + 3) / 4 * 4; }
}  // namespace

LearnedPrecond::LearnedPrecond(Context &ctx, const icarat_learned_opts &o) : ctx_(ctx), o_(o) {
  ctx_.ensure_device();
  if (o_.base != 0 && o_.base != 1) throw Error(kInvalidArgument, "learned base must be 0 or 1");
  if (!(o_.correction_scale > 0.f)) o_.correction_scale = 1.f;
  mg_base_ = o_.base == 1;
  if (mg_base_ && !ctx_.mg_) ctx_.mg_setup(mg_default_opts());
  const Lattice &L = ctx_.lattice();
  g0_ = NetGrid{round4((L.nx + 1) / 2), round4((L.ny + 1) / 2), round4((L.nz + 1) / 2)};
  ncell_ = g0_.nvox();
  hipStream_t s = ctx_.stream();
  net_.reset(new UNet(g0_, s, o_.seed));
  feat_.alloc(ncell_);
  s3_.alloc(ncell_ * 3);
  out_.alloc(ncell_ * 8);
  xin_.alloc(ncell_ * 8);
  scale_.alloc(4);
  z_.alloc(ctx_.n3_);
  corr_.alloc(ctx_.n3_);
  ones_.alloc(ctx_.n3_);
  HIPCHK(hipMemsetAsync(z_.p, 0, z_.bytes(), s));
  HIPCHK(hipMemsetAsync(scale_.p, 0, scale_.bytes(), s));
  {
    // 1 at free DOFs, 0 elsewhere: the K.p kernel then forms p = z + beta p_old
    std::vector<double> ones(ctx_.n3_, 0.0);
    const auto &fm = ctx_.problem().free_mask();
    for (long long n = 0; n < L.numnp; ++n)
      for (int c = 0; c < 3; ++c) ones[c * L.Np + n] = ((fm[n] >> c) & 1u) ? 1.0 : 0.0;
    HIPCHK(hipMemcpyAsync(ones_.p, ones.data(), ones_.bytes(), hipMemcpyHostToDevice, s));
    ctx_.sync();
  }
  cap_ = std::max(1, o_.max_samples);
  valid_.assign(cap_, 0);
  sx_.alloc((size_t)cap_ * ncell_ * 8);
  szr_.alloc((size_t)cap_ * ncell_ * 3);
  st_.alloc((size_t)cap_ * ncell_ * 3);
  ssig_.alloc(cap_);
  rxs_.alloc(ncell_ * 3);
  rng_ = o_.seed * 0x9e3779b97f4a7c15ULL + 1;
}

LearnedPrecond::~LearnedPrecond() { drop_graphs(); }

// ---- harvesting -------------------------------------------------------------
void LearnedPrecond::begin_harvest() {
  harvesting_ = true;
  harvest_batches_ = 0;
  pending_.clear();
  launch_density_feature(ctx_.lattice(), g0_, ctx_.d_young, feat_, ctx_.stream());
}

const double *LearnedPrecond::base_vec(const double *r, const double **dinv) {
  if (!mg_base_) {
    *dinv = ctx_.d_dinv.p;
    return r;
  }
  *dinv = ones_.p;
  return ctx_.mg_->vcycle(r);
}

void LearnedPrecond::snapshot() {
  if (!harvesting_) return;
  // iterations per batch of the harvesting solve (Jacobi-PCG or multigrid PCG)
  const int per = mg_base_ ? ctx_.mg_->batch_iters() : Context::kBatch;
  if (harvest_batches_++ % std::max(1, o_.snapshot_every / per) != 0) return;
  // the snapshots of one solve never overwrite each other
  if ((int)pending_.size() >= cap_) return;
  hipStream_t s = ctx_.stream();
  const Lattice &L = ctx_.lattice();
  const int k = next_;
  next_ = (next_ + 1) % cap_;
  valid_[k] = 0;
  // B(r_k): D^-1 r_k, or the V-cycle the multigrid-PCG iteration just applied
  // to r_k (its result buffer holds V(r) at a batch boundary)
  const double *bz = mg_base_ ? ctx_.mg_->result() : ctx_.d_r.p;
  launch_restrict_residual(L, g0_, bz, mg_base_ ? ones_.p : ctx_.d_dinv.p, s3_, ctx_.d_partials,
                           scale_, s);
  launch_pack_input(g0_, feat_, s3_, scale_, sx_.p + (size_t)k * ncell_ * 8,
                    szr_.p + (size_t)k * ncell_ * 3, s);
  HIPCHK(hipMemcpyAsync(ssig_.p + k, scale_.p, sizeof(double), hipMemcpyDeviceToDevice, s));
  launch_restrict_vec(L, g0_, ctx_.d_x, st_.p + (size_t)k * ncell_ * 3, s);
  pending_.push_back(k);
}

void LearnedPrecond::end_harvest(bool converged) {
  if (!harvesting_) return;
  harvesting_ = false;
  ++harvested_;
  hipStream_t s = ctx_.stream();
  if (converged) {
    launch_restrict_vec(ctx_.lattice(), g0_, ctx_.d_x, rxs_, s);
    std::vector<double> sig(cap_);
    HIPCHK(hipMemcpyAsync(sig.data(), ssig_.p, cap_ * sizeof(double), hipMemcpyDeviceToHost, s));
    ctx_.sync();
    std::vector<int> made;
    for (int k : pending_) {
      // a zero residual carries no information
      if (!(sig[k] > 0.0)) continue;
      float *t = st_.p + (size_t)k * ncell_ * 3;
      launch_make_target(ncell_ * 3, rxs_, t, 1.0 / sig[k], szr_.p + (size_t)k * ncell_ * 3, t, s);
      made.push_back(k);
    }
    if (!made.empty() && !gain_set_) {
      // output scale: the median RMS of the first solve's targets, so that the
      // network works at unit scale
      std::vector<float> h((size_t)ncell_ * 3);
      std::vector<double> rms;
      for (int k : made) {
        HIPCHK(hipMemcpyAsync(h.data(), st_.p + (size_t)k * ncell_ * 3, h.size() * sizeof(float),
                              hipMemcpyDeviceToHost, s));
        ctx_.sync();
        double a = 0.0;
        for (float v : h) a += (double)v * v;
        rms.push_back(std::sqrt(a / (double)h.size()));
      }
      std::nth_element(rms.begin(), rms.begin() + rms.size() / 2, rms.end());
      gain_ = rms[rms.size() / 2] > 0.0 ? rms[rms.size() / 2] : 1.0;
      gain_set_ = true;
    }
    for (int k : made) {
      launch_scale_f32(ncell_ * 3, (float)(1.0 / gain_), st_.p + (size_t)k * ncell_ * 3, s);
      valid_[k] = 1;
    }
    ctx_.sync();
  }
  pending_.clear();
}

void LearnedPrecond::set_params(const float *host) {
  net_->set_params(host);
  active_ = true;
}

double LearnedPrecond::evaluate() {
  hipStream_t s = ctx_.stream();
  double acc = 0.0;
  int cnt = 0;
  for (int k = 0; k < cap_; ++k) {
    if (!valid_[k]) continue;
    net_->eval_loss(sx_.p + (size_t)k * ncell_ * 8, st_.p + (size_t)k * ncell_ * 3);
    double l2[2];
    HIPCHK(hipMemcpyAsync(l2, net_->loss2(), 2 * sizeof(double), hipMemcpyDeviceToHost, s));
    ctx_.sync();
    if (l2[1] > 0.0) {
      acc += l2[0] / l2[1];
      ++cnt;
    }
  }
  return cnt ? acc / cnt : -1.0;
}

int LearnedPrecond::samples() const {
  int n = 0;
  for (char v : valid_) n += v ? 1 : 0;
  return n;
}

double LearnedPrecond::train(int steps) {
  std::vector<int> slots;
  for (int k = 0; k < cap_; ++k)
    if (valid_[k]) slots.push_back(k);
  if (slots.empty() || steps <= 0) return -1.0;
  hipStream_t s = ctx_.stream();
  double acc = 0.0;
  int cnt = 0;
  const int tail = std::max(1, std::min(steps, 20));
  for (int i = 0; i < steps; ++i) {
    rng_ ^= rng_ << 13;
    rng_ ^= rng_ >> 7;
    rng_ ^= rng_ << 17;
    const int k = slots[rng_ % (uint64_t)slots.size()];
    net_->train_step(sx_.p + (size_t)k * ncell_ * 8, st_.p + (size_t)k * ncell_ * 3, o_.lr);
    if (i >= steps - tail) {
      double l2[2];
      HIPCHK(hipMemcpyAsync(l2, net_->loss2(), 2 * sizeof(double), hipMemcpyDeviceToHost, s));
      ctx_.sync();
      if (l2[1] > 0.0) {
        acc += l2[0] / l2[1];
        ++cnt;
      }
    }
  }
  total_steps_ += steps;
  last_loss_ = cnt ? acc / cnt : -1.0;
  // the gate: the network is used only if it explains enough of the targets
  val_loss_ = evaluate();
  active_ = val_loss_ >= 0.0 && val_loss_ < (double)o_.gate;
  return last_loss_;
}

// ---- preconditioning and flexible PCG -----------------------------------------
void LearnedPrecond::precondition(const double *r) {
  hipStream_t s = ctx_.stream();
  const Lattice &L = ctx_.lattice();
  const double *dinv = nullptr;
  const double *b = base_vec(r, &dinv);
  if (!active_) {
    launch_prolong_add(L, g0_, out_, 8, scale_, b, dinv, ctx_.d_free, z_, 0, 0.0, s);
    return;
  }
  launch_restrict_residual(L, g0_, b, dinv, s3_, ctx_.d_partials, scale_, s);
  launch_pack_input(g0_, feat_, s3_, scale_, xin_, nullptr, s);
  net_->forward(xin_, out_);
  launch_prolong_add(L, g0_, out_, 8, scale_, b, dinv, ctx_.d_free, z_, 1,
                     gain_ * (double)o_.correction_scale, s, corr_.p);
  // z.r > 0 whatever the network returns: the correction is shrunk where it
  // would cancel more than half of the base step's descent
  launch_descent_guard(ctx_.n3_, b, dinv, corr_, r, z_, ctx_.d_partials, scale_, s);
}

void LearnedPrecond::apply(const double *r, double *z) {
  if (mg_base_) ctx_.mg_->refresh_now();
  launch_density_feature(ctx_.lattice(), g0_, ctx_.d_young, feat_, ctx_.stream());
  precondition(r);
  HIPCHK(hipMemcpyAsync(z, z_.p, z_.bytes(), hipMemcpyDeviceToDevice, ctx_.stream()));
  ctx_.sync();
}

void LearnedPrecond::enqueue_batch(int kmode, int par) {
  hipStream_t s = ctx_.stream();
  Context &c = ctx_;
  const Lattice &L = c.lattice();
  for (int j = 0; j < std::max(1, o_.batch_iters); ++j) {
    double *p_old = c.d_p[(j + par) & 1], *p_new = c.d_p[(j + par + 1) & 1];
    // p = z + beta p_old (the "dinv" operand is 1 at free DOFs), q = K p, alpha
    if (kmode < 0)
      launch_pcg_stencil64(L, c.kc_, c.d_young, c.d_free, z_, ones_, p_old, p_new, c.d_q, c.d_pcg,
                           c.d_partials, c.d_counter, s);
    else
      launch_pcg_stencil_mfma64(L, kmode, c.d_ke32, c.d_young32, c.d_free, z_, ones_, p_old, p_new,
                                c.d_q, c.d_pcg, c.d_partials, s);
    launch_fpcg_update(c.n3_, c.d_x, c.d_r, p_new, c.d_q, c.d_pcg, c.d_partials, s);
    precondition(c.d_r);
    launch_zdot_beta(c.n3_, z_, c.d_r, c.d_q, c.d_pcg, c.d_partials, s);
  }
}

void LearnedPrecond::drop_graphs() {
  for (int m = 0; m < 3; ++m)
    for (int p = 0; p < 2; ++p) {
      if (exec_[m][p]) (void)hipGraphExecDestroy(exec_[m][p]);
      if (graph_[m][p]) (void)hipGraphDestroy(graph_[m][p]);
      exec_[m][p] = nullptr;
      graph_[m][p] = nullptr;
    }
}

void LearnedPrecond::build_graph(int kmode) {
  const int m = kmode + 1;
  // the captured batch bakes in whether the network runs and its gain
  if (graph_active_ != active_ || graph_gain_ != gain_ * (double)o_.correction_scale) {
    drop_graphs();
    graph_active_ = active_;
    graph_gain_ = gain_ * (double)o_.correction_scale;
  }
  if (!ctx_.use_graphs_ || exec_[m][0]) return;
  hipStream_t s = ctx_.stream();
  for (int par = 0; par < 2; ++par) {
    HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    try {
      enqueue_batch(kmode, par);
    } catch (...) {
      hipGraph_t g = nullptr;
      (void)hipStreamEndCapture(s, &g);
      if (g) (void)hipGraphDestroy(g);
      throw;
    }
    HIPCHK(hipStreamEndCapture(s, &graph_[m][par]));
    HIPCHK(hipGraphInstantiate(&exec_[m][par], graph_[m][par], nullptr, nullptr, 0));
  }
}

void LearnedPrecond::solve(const icarat_solve_opts &o, icarat_solve_stats &st) {
  Context &c = ctx_;
  hipStream_t s = c.stream();
  const Lattice &L = c.lattice();
  const int kmode = mg_base_ ? -1
                             : (o.precision == ICARAT_HIP_FP64 ? -1
                                                               : (o.precision == ICARAT_HIP_FP32 ? 0 : 1));
  if (mg_base_) {
    if (!active_) {
      // without the network the preconditioner is the V-cycle: multigrid PCG
      ctx_.mg_->solve(o, st);
      return;
    }
    if (o.precision == ICARAT_HIP_BF16)
      throw Error(kInvalidArgument, "the multigrid base runs an FP64 or an FP32 fine level");
    ctx_.mg_->set_fine_fp32(o.precision == ICARAT_HIP_FP32);
    ctx_.mg_->refresh_now();
  }
  if (!active_) {
    // without the network M = D^-1 is fixed: plain Jacobi-PCG (whose beta
    // keeps the short recurrence exact) instead of the flexible variant
    const PcgScalars r = c.pcg64(o.tol, o.maxit, false, kmode, kmode < 0 ? 0.0 : 0.1);
    if (r.status == 2) throw Error(kNumerical, "PCG breakdown");
    st.iters = r.iter;
    st.refine_steps = c.replacements_;
    st.fallback = 0;
    st.rec_relres = r.bnorm2 > 0.0 ? std::sqrt(r.rr) / std::sqrt(r.bnorm2) : 0.0;
    st.status = r.status == 1 ? ICARAT_HIP_NOT_CONVERGED : ICARAT_HIP_OK;
    return;
  }
  build_graph(kmode);
  launch_density_feature(L, g0_, c.d_young, feat_, s);
  PcgScalars *h = c.h_pcg.p;
  PcgScalars s0{};
  s0.tol = o.tol;
  s0.maxit = o.maxit;
  *h = s0;
  HIPCHK(hipMemcpyAsync(c.d_pcg.p, h, sizeof(PcgScalars), hipMemcpyHostToDevice, s));
  // x = 0, r = b; then z = M(r), rho = z.r, beta = 0
  launch_pcg_init64(c.n3_, c.d_b, c.d_x, c.d_r, c.d_p[0], c.d_dinv, c.d_pcg, c.d_partials,
                    c.d_counter, s);
  HIPCHK(hipMemcpyAsync(h, c.d_pcg.p, sizeof(PcgScalars), hipMemcpyDeviceToHost, s));
  c.sync();
  auto restart = [&]() {
    h->restart = 1;
    h->alpha = 0.0;
    HIPCHK(hipMemcpyAsync(c.d_pcg.p, h, sizeof(PcgScalars), hipMemcpyHostToDevice, s));
    precondition(c.d_r);
    launch_zdot_beta(c.n3_, z_, c.d_r, c.d_q, c.d_pcg, c.d_partials, s);
    HIPCHK(hipMemcpyAsync(h, c.d_pcg.p, sizeof(PcgScalars), hipMemcpyDeviceToHost, s));
    c.sync();
  };
  if (!h->done) restart();
  const double bnorm = std::sqrt(h->bnorm2);
  const int batch = std::max(1, o_.batch_iters);
  double best = std::sqrt(h->rr);
  int best_iter = 0, repl = 0;
  bool fail = false;
  while (!h->done || h->status == 0) {
    if (!h->done) {
      const long long max_batches = (long long)(o.maxit - h->iter) / batch + 3;
      for (long long b = 0; b < max_batches && !h->done; ++b) {
        // the p buffers alternate each iteration; a batch starts at the parity of iter
        if (ctx_.use_graphs_)
          HIPCHK(hipGraphLaunch(exec_[kmode + 1][h->iter & 1], s));
        else
          enqueue_batch(kmode, h->iter & 1);
        HIPCHK(hipMemcpyAsync(h, c.d_pcg.p, sizeof(PcgScalars), hipMemcpyDeviceToHost, s));
        c.sync();
        const double rn = std::sqrt(h->rr);
        if (rn < 0.5 * best) {
          best = rn;
          best_iter = h->iter;
        } else if (h->iter - best_iter > o_.stall_iters) {
          fail = true;  // no halving of the residual for stall_iters iterations
          break;
        }
      }
      if (fail) break;
      if (!h->done) throw Error(kInternal, "flexible PCG did not reach its stop test");
    }
    if (h->status != 0) break;
    // verified stop: the FP64 true residual replaces the recursive one
    launch_apply_k64(L, c.kc_, c.d_young, c.d_free, c.d_x, c.d_y, s);
    launch_replace64(c.n3_, c.d_b, c.d_y, c.d_r, c.d_dinv, c.d_pcg, c.d_partials, s);
    HIPCHK(hipMemcpyAsync(h, c.d_pcg.p, sizeof(PcgScalars), hipMemcpyDeviceToHost, s));
    c.sync();
    if (h->done) break;
    if (++repl > 20) {
      fail = true;
      break;
    }
    restart();
  }
  int iters = h->iter;
  st.fallback = 0;
  st.refine_steps = repl;
  const bool converged = h->done && h->status == 0;
  if (!converged && (fail || h->status == 2 || h->status == 4) && iters < o.maxit) {
    // why: 1 the residual stalled (or too many verifications), 2 a breakdown,
    // 3 the preconditioned residual was not a descent direction (z.r <= 0)
    st.fallback = fail ? 1 : (h->status == 4 ? 3 : 2);
    fallback_iter_ = iters;
    if (mg_base_) {
      // multigrid PCG from the current iterate
      icarat_solve_opts fo = o;
      fo.maxit = o.maxit - iters;
      icarat_solve_stats fs{};
      ctx_.mg_->solve(fo, fs, true);
      iters += fs.iters;
      st.iters = iters;
      st.refine_steps = repl + fs.refine_steps;
      st.rec_relres = fs.rec_relres;
      st.status = fs.status;
      return;
    }
    // Jacobi-PCG from the current iterate
    const PcgScalars r = c.pcg64(o.tol, o.maxit - iters, true, kmode, kmode < 0 ? 0.0 : 0.1);
    iters += r.iter;
    if (r.status == 2) throw Error(kNumerical, "PCG breakdown");
    *h = r;
  }
  st.iters = iters;
  st.rec_relres = bnorm > 0.0 ? std::sqrt(h->rr) / bnorm : 0.0;
  st.status = (h->done && h->status == 0) || bnorm == 0.0 ? ICARAT_HIP_OK : ICARAT_HIP_NOT_CONVERGED;
}

}  // namespace icarat_hip
