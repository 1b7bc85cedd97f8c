// learned_kernels.hpp -- launchers of the lattice <-> cell-grid transfers and
// the flexible-PCG reductions of the learned preconditioner (k_learned.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "device_common.hpp"
#include "hip_util.hpp"
#include "unet_kernels.hpp"

namespace icarat_hip {

// s3 = R(dinv r) per cell; scale[0] = sigma (RMS of s3), scale[1] = 1/sigma,
// scale[2] scratch
void launch_restrict_residual(const Lattice &L, const NetGrid &g, const double *r, const double *dinv,
                              float *s3, double *partials, double *scale, hipStream_t s);
void launch_restrict_vec(const Lattice &L, const NetGrid &g, const double *x, float *out3,
                         hipStream_t s);
void launch_pack_input(const NetGrid &g, const float *feat, const float *s3, const double *scale,
                       __bf16 *xin, float *zr, hipStream_t s);
void launch_density_feature(const Lattice &L, const NetGrid &g, const double *young, float *feat,
                            hipStream_t s);
// z = dinv r + gain sigma P(out) (use_net = 0: z = dinv r); corr_out (may be
// null) receives the correction gain sigma P(out) alone
void launch_prolong_add(const Lattice &L, const NetGrid &g, const float *out, int cs,
                        const double *scale, const double *r, const double *dinv,
                        const uint8_t *fmask, double *z, int use_net, double gain, hipStream_t s,
                        double *corr_out = nullptr);
// Descent guard of z = dinv base + c against the residual r: the correction is
// shrunk to gamma c (gamma in scale[3]) so that z.r >= (dinv base).r / 2 > 0
// (gamma = 1 leaves z as it is)
void launch_descent_guard(long long n, const double *base, const double *dinv, const double *c,
                          const double *r, double *z, double *partials, double *scale,
                          hipStream_t s);
void launch_scale_f32(long long n, float a, float *x, hipStream_t s);
void launch_make_target(long long n, const float *rxs, const float *rxk, double inv_sigma,
                        const float *zr, float *t, hipStream_t s);
// rho = z.r, beta = -alpha z.q / rho_old (Polak-Ribiere; 0 after a restart);
// status 4 when z.r <= 0
void launch_zdot_beta(long long n, const double *z, const double *r, const double *q, PcgScalars *S,
                      double *partials, hipStream_t s);
// x += alpha p, r -= alpha q, r.r and the stop tests
void launch_fpcg_update(long long n, double *x, double *r, const double *p, const double *q,
                        PcgScalars *S, double *partials, hipStream_t s);
// Flexible PCG with the scalar steps folded into neighbouring kernels (used by
// the multigrid solve).  launch_fpcg_update_alpha: alpha = rho / p.q from the
// stencil's partials (Gs blocks) in every block, x and r updated, r.r partials
// to part_up, and with z0 != nullptr also z0 = omega dinv r.  With xlo !=
// nullptr, x is updated as a compensated sum whose low part is xlo.  Returns
// the partials count (per load case: nc cases stacked with stride cs, their
// scalars S[c], the stencil's partials at c x Gs, ours at c x the count).
unsigned launch_fpcg_update_alpha(long long n, double *x, double *r, const double *p,
                                  const double *q, PcgScalars *S, const double *part_st,
                                  unsigned Gs, double *part_up, const double *dinv, double omega,
                                  double *z0, hipStream_t s, double *xlo = nullptr, int nc = 1,
                                  long long cs = 0);
// the same with z0 in FP32 (the FP32 fine level of the multigrid V-cycle)
unsigned launch_fpcg_update_alpha(long long n, double *x, double *r, const double *p,
                                  const double *q, PcgScalars *S, const double *part_st,
                                  unsigned Gs, double *part_up, const double *dinv, double omega,
                                  float *z0, hipStream_t s, double *xlo = nullptr);
// z.r and z.q partials; returns their count per quantity (nc load cases:
// vectors stacked with stride cs, case c's partials at 2 c x the count)
unsigned launch_zdot_partials(long long n, const double *z, const double *r, const double *q,
                              double *part_z, hipStream_t s, int nc = 1, long long cs = 0);
// the next beta step restarts (beta = 0) unless the solve is done
void launch_fpcg_restart_flag(PcgScalars *S, hipStream_t s);
// stop tests from the r.r partials, then rho' = z.r and beta (Polak-Ribiere);
// nc load cases: one block per case, its partials at c x Gu and 2 c x Gz, S[c]
void launch_fpcg_check_beta(const double *part_up, unsigned Gu, const double *part_z, unsigned Gz,
                            PcgScalars *S, hipStream_t s, int nc = 1);

}  // namespace icarat_hip
