// kernels.hpp -- host-side launchers of the HIP kernels (one declaration per
// kernel family; the definitions live next to their kernels in k_*.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "device_common.hpp"
#include "hip_util.hpp"

namespace icarat_hip {

// ---- k_fp64.hip ----------------------------------------------------------
// `ke` below is the 576-entry KE0 in device memory (row-major 24x24).

// KE0 as the FP64 stencil kernels take it.  With sym > 0 the kernels use the
// orbit values v (kernel arguments, resident in scalar registers) through the
// compile-time orbit tables of ke_sym_table.hpp instead of loading KE0:
// sym 1 = box symmetry (48 values), sym 2 = cube symmetry (12 values).
template <class T>
struct KeValsT {
  T v[48];
};
using KeVals = KeValsT<double>;
struct KeCoef {
  const double *ke = nullptr;  // device KE0
  int sym = 0;
  int kernel = 0;              // 0: z-marching column kernel, 1: 3-D tile kernel
  int nchunks = 0;             // z-marching: z-chunks per column (0: default, capped by it)
  int xlo = 0;                 // 1: node column x = 0 holds no free DOF (the z-marching
                               // tiles start at x = 1)
  KeVals vals{};
  // the same in FP32 for the FP32 z-marching K.u (multigrid fine level)
  const float *ke32 = nullptr; // device KE0 in FP32
  KeValsT<float> vals32{};
};
// Orbit values of a host KE0 (sym = 2 if the cube table reproduces it to
// 1e-13 of max |KE0|, else 1 for the box table, else 0).
KeCoef make_ke_coef(const double *ke_host, const double *ke_dev);
// as make_ke_coef restricted to one class (sym = cls if it reproduces KE0, else 0)
KeCoef make_ke_coef_class(const double *ke_host, const double *ke_dev, int cls);
// young[e] = SIMP(x[e]) (reference problem/elasticity.hpp:13-17)
void launch_young(const double *x, double *young, long long nelx, double y1, double y2, double p,
                  hipStream_t s);
// dinv[c*Np+n] = 1 / diag K at free DOFs, 0 elsewhere
void launch_diag_inv(const Lattice &L, const double *ke, const double *young,
                     const uint8_t *freemask, double *dinv, double *diag, hipStream_t s);
// number of workgroups of the largest K.u launch (sizes the partials buffer)
long long stencil_blocks(const Lattice &L);
// default z-chunk count of the z-marching kernel (one round of resident workgroups)
int stencil_nchunks(const Lattice &L);
// y = K u on the lattice; y masked to free DOFs.  u is used as given
// (Dirichlet values included).  ke.kernel selects the z-marching or the 3-D
// halo-tile kernel.
void launch_apply_k64(const Lattice &L, const KeCoef &ke, const double *young,
                      const uint8_t *freemask, const double *u, double *y, hipStream_t s);
// FP32 z-marching K.u (FP32 node values, young32 and KE0 in FP32, FP32
// accumulation): y = K x masked to free DOFs, stored in FP64; the _jacobi form is the fused
// smoother sweep of launch_apply_k64_jacobi with an FP32 x and the FP64
// epilogue (xout, b, dinv, q and the partials in FP64).  Either ignores
// ke.kernel (always the z-marching shape).
void launch_apply_k32(const Lattice &L, const KeCoef &ke, const float *young32,
                      const uint8_t *freemask, const float *x, double *y, hipStream_t s);
unsigned launch_apply_k32_jacobi(const Lattice &L, const KeCoef &ke, const float *young32,
                                 const uint8_t *freemask, const float *x, double *xout,
                                 const double *b, const double *dinv, double omega,
                                 const double *q, double *partials, hipStream_t s);
// Arguments of the fused epilogues of the z-marching K.u: the smoother sweep
// (y = x + omega dinv (b - K x), with the partials of y.b and y.q when q is
// given), or with u given the start of a W-cycle's second coarse visit
// (u = x, y = b - K x masked, and with x0 also x0 = omega dinv y).
struct JacEpilogue {
  const double *b = nullptr;
  const double *q = nullptr;
  double omega = 0.0;
  double *u = nullptr;
  double *x0 = nullptr;
};
// The multigrid smoother's sweep with K.x in one z-marching pass:
// xout = x + omega dinv (b - K x) (x unchanged) and, with q and partials given,
// the block partials of xout.b and xout.q (2 per block, the layout of
// store_partials); returns the block count.  Needs the z-marching kernel shape.
unsigned launch_apply_k64_jacobi(const Lattice &L, const KeCoef &ke, const double *young,
                                 const uint8_t *freemask, const double *x, double *xout,
                                 const double *b, const double *dinv, double omega,
                                 const double *q, double *partials, hipStream_t s);
// Load cases of the multigrid solve in one launch (case c's vectors at c x cs,
// its scalars S[c], its partials -- one or two per block -- at c x (one or two)
// x the returned count, the layout of the single-case launch).
// launch_pcg_stencil64_mcz: p_new = z + beta_c p_old, q = K p_new and the p.q
// partials.  launch_apply_k64_mc: y = K x (b == nullptr); with partials, the
// smoother sweep y = x + omega dinv (b - K x) with the y.b and y.q partials;
// without, the sweep alone, or with u the W-cycle residual of
// launch_apply_k64_wres (u = x, y = b - K x, x0 = omega dinv y).  Both need
// the z-marching kernel shape.
unsigned launch_pcg_stencil64_mcz(const Lattice &L, const KeCoef &ke, const double *young,
                                  const uint8_t *freemask, const double *z, const double *p_old,
                                  double *p_new, double *q, PcgScalars *S, double *partials,
                                  int nc, long long cs, hipStream_t s);
unsigned launch_apply_k64_mc(const Lattice &L, const KeCoef &ke, const double *young,
                             const uint8_t *freemask, const double *x, double *y, const double *b,
                             const double *dinv, double omega, const double *q, double *partials,
                             int nc, long long cs, hipStream_t s, double *u = nullptr,
                             double *x0 = nullptr);
// The start of a W-cycle's second visit of a z-marching level in one pass:
// u = xc, r = b - K xc (masked) and, with x0, x0 = omega dinv r (xc may not
// alias u, r or x0).  Needs the z-marching kernel shape.
void launch_apply_k64_wres(const Lattice &L, const KeCoef &ke, const double *young,
                           const uint8_t *freemask, const double *xc, const double *b,
                           const double *dinv, double omega, double *u, double *r, double *x0,
                           hipStream_t s);
// One fused Jacobi-PCG iteration, stencil half: p_new = dinv*r + beta*p_old
// (dinv == nullptr: p_new = r + beta*p_old, for a preconditioned r);
// q = K p_new; p.q partials (returns their count); with finalize, alpha = rho /
// p.q by a finalize kernel.
unsigned launch_pcg_stencil64(const Lattice &L, const KeCoef &ke, const double *young,
                              const uint8_t *freemask, const double *r, const double *dinv,
                              const double *p_old, double *p_new, double *q, PcgScalars *S,
                              double *partials, unsigned *counter, hipStream_t s,
                              bool finalize = true);
// Fused batches: the alpha and beta finalizes run inside the next kernel (every
// block reduces the partials itself).  In-batch iteration j: the stencil (j >= 1
// reduces partials_up of update j-1 over Gu blocks for beta and the stop test;
// j == 0 takes beta from the scalars) writes partials_st and returns its block
// count; the update reduces partials_st over Gs blocks for alpha and writes
// partials_up (2 per block) and returns its block count.  A batch ends with
// launch_finalize(kFinPcgBeta, 2, partials_up, Gu, ...).
unsigned launch_pcg_stencil64_fused(const Lattice &L, const KeCoef &ke, const double *young,
                                    const uint8_t *freemask, const double *r, const double *dinv,
                                    const double *p_old, double *p_new, double *q, PcgScalars *S,
                                    double *partials_st, const double *partials_up, unsigned Gu,
                                    int j, hipStream_t s);
unsigned launch_pcg_update64_fused(long long n, double *x, double *r, const double *p,
                                   const double *q, const double *dinv, PcgScalars *S,
                                   double *partials_up, const double *partials_st, unsigned Gs,
                                   int j, hipStream_t s);
// Jacobi-PCG batch steps of nc load cases in one launch each (k_stencil64z_mc,
// k_pcg_update64f_mc): case c's r, q, x at c*cs, its two p buffers at c*2*cs,
// scalars S[c]; stencil partials nt per case, update partials 2*UPD per case.
// Each case's arithmetic and sums are those of its own single-case batch.
// y = K x on the matrix cores (v_mfma_f64_4x4x4_4b element contraction, node
// gather); the z-marching tile and z-chunks of launch_apply_k64
void launch_apply_k64_mfma(const Lattice &L, const KeCoef &ke, const double *young,
                           const uint8_t *freemask, const double *x, double *y, hipStream_t s);
unsigned pcg_update_blocks();  // blocks of the fused update kernels (partials per quantity)
unsigned launch_pcg_stencil64_mc(const Lattice &L, const KeCoef &ke, const double *young,
                                 const uint8_t *freemask, const double *r, const double *dinv,
                                 double *p, double *q, PcgScalars *S, double *partials_st,
                                 const double *partials_up, unsigned Gu, int j, int nc,
                                 long long cs, hipStream_t s);
unsigned launch_pcg_update64_mc(long long n, double *x, double *r, const double *p,
                                const double *q, const double *dinv, PcgScalars *S,
                                double *partials_up, const double *partials_st, unsigned Gs,
                                int j, int nc, long long cs, hipStream_t s);
// Vector half: x += alpha p; r -= alpha q; rho' = r.dinv.r; r.r; beta; stop test.
void launch_pcg_update64(long long n, double *x, double *r, const double *p, const double *q,
                         const double *dinv, PcgScalars *S, double *partials, unsigned *counter,
                         hipStream_t s, bool finalize = true);
// x = 0, r = b, p = 0, rho = b.dinv.b, bnorm2 = b.b
void launch_pcg_init64(long long n, const double *b, double *x, double *r, double *p0,
                       const double *dinv, PcgScalars *S, double *partials, unsigned *counter,
                       hipStream_t s);
// reliable update / verified stop: r = b - y with y = K x; rho and r.r of
// the new residual through the kFinReplace finalize
void launch_replace64(long long n, const double *b, const double *y, double *r,
                      const double *dinv, PcgScalars *S, double *partials, hipStream_t s);
// res = b - y on free DOFs; sums ||res||^2 and ||b||^2 into out2[0..1]
void launch_residual_norms(long long n, const double *b, const double *y, double *out2,
                           double *partials, unsigned *counter, hipStream_t s);
// out = a . b  (length n)
void launch_dot(long long n, const double *a, const double *b, double *out, double *partials,
                unsigned *counter, hipStream_t s);
// nodal values with |v| < 1e-15 flushed (reference analysis/analysis_utility.hpp:26-28)
void launch_flush_small(long long n, const double *u, double *v, hipStream_t s);
// dfds[e] = -(E'/E) u_e^T (E KE0) u_e, dgds[e] = vol_e / vnorm
void launch_sensitivity64(const Lattice &L, const double *ke, const double *dens,
                          const double *nodeval, double y1, double y2, double p, double vol_e,
                          double vnorm, double *dfds, double *dgds, hipStream_t s);
// element-mean stress (stress may be null; nelx x 6 Voigt) and von Mises stress
// of the nodal values, voxel edge lengths h (Elastic::makeMisesStress)
void launch_mises64(const Lattice &L, const double *dens, const double *nodeval, double y1,
                    double y2, double p, double nu, const double h[3], double *mises,
                    double *stress, hipStream_t s);
// lattice <-> reference numeq ordering (dof[3*n+c] is the reference DOF number)
void launch_numeq_to_lattice(const Lattice &L, const int *dof, int numeq, const double *v,
                             const double *dirval, double *lat, hipStream_t s);
void launch_lattice_to_numeq(const Lattice &L, const int *dof, int numeq, const double *lat,
                             double *v, hipStream_t s);
void launch_axpby(long long n, double a, const double *x, double b, double *y, hipStream_t s);
// y = x (plain copy; the calibration stream for bandwidth and traffic counters)
void launch_copy64(long long n, const double *x, double *y, hipStream_t s);

// ---- k_filter.hip ---------------------------------------------------------
// Filter neighbourhood in device memory: lin[t] is the LDS offset of entry t
// for a halo of width R, w[t] = rmin - distance.
struct FilterTable {
  int count;          // number of offsets
  int R;              // halo width in elements (1..3)
  const int *lin;     // device
  const double *w;    // device
};
// LDS linear offsets of (dx,dy,dz) for the tile shape used with halo R
void filter_lds_offsets(const int *off, int count, int R, int *lin);
// is_sens == 0: out = (sum w x)/(sum w), aa = sum w
// is_sens == 1: out = sum w v / aa_j
void launch_density_filter(const Lattice &L, const FilterTable &T, const double *in,
                           double *aa, double *out, int is_sens, hipStream_t s);
// Sigmund's sensitivity filter: out_i = (x_i V)/(x_i sum_j w_ij) *
// sum_j (w_ij x_j dfds_j)/(x_j V), neighbours in ascending element id
void launch_sensitivity_filter(const Lattice &L, const FilterTable &T, const double *x,
                               const double *dfds, double *out, double vol_e, hipStream_t s);
void launch_threshold(long long n, const double *x, double beta, double T, double *out,
                      int deriv, hipStream_t s);
void launch_mul(long long n, const double *a, double *b, hipStream_t s);
// One OC bisection step (x' at lambda, filter, optional projection, volume
// sum, bisection update of the device state).
void launch_oc_step(const Lattice &L, const FilterTable &T, const double *x, const double *dc,
                    const double *dv, double xmin, double xmax,
                    int use_projection, double beta, double vol_e, OcState *S,
                    double *partials, unsigned *counter, hipStream_t s);
void launch_oc_final(long long n, double *x, const double *dc, const double *dv, double xmin,
                     double xmax, const OcState *S, hipStream_t s);
// The OC bisection step in pieces, for a filter that is not a fused stencil
// (the Helmholtz filter): out = x' at the current lambda; then, after the
// caller has filtered out, the volume of the (projected) filtered x' and the
// bisection update of the device state.  Both skip once the bisection is done.
void launch_oc_candidate(long long n, const double *x, const double *dc, const double *dv,
                         double xmin, double xmax, const OcState *S, double *out, hipStream_t s);
void launch_oc_volume(long long n, const double *xf, int use_projection, double beta, double vol_e,
                      OcState *S, double *partials, hipStream_t s);
// sum of x (or of the projection of x) times vol_e, into *out
void launch_volume(long long n, const double *x, double vol_e, double *out, double *partials,
                   unsigned *counter, hipStream_t s);

}  // namespace icarat_hip
