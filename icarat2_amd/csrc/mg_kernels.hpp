// mg_kernels.hpp -- launchers of the multigrid kernels (k_mg.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "device_common.hpp"
#include "hip_util.hpp"
#include "kernels.hpp"

namespace icarat_hip {

// bc = R rf on the coarse lattice (transpose of trilinear prolongation), masked;
// with yf: bc = R (rf - yf) (the fine residual formed on the fly); with dinv_c:
// also xc = omega dinv_c bc (the first Jacobi sweep from zero)
void launch_mg_restrict(const Lattice &Lf, const Lattice &Lc, const double *rf, const double *yf,
                        const uint8_t *fmask_c, double *bc, const double *dinv_c, double omega,
                        double *xc, hipStream_t s, int nc = 1, long long sf = 0,
                        long long sc = 0);
// xf += P xc (trilinear interpolation), masked; add = false: xf = P xc at the
// free DOFs (constrained DOFs keep their value)
void launch_mg_prolong_add(const Lattice &Lf, const Lattice &Lc, const double *xc,
                           const uint8_t *fmask_f, double *xf, hipStream_t s, bool add = true,
                           const double *xc2 = nullptr, int nc = 1, long long sf = 0,
                           long long sc = 0);
// FP32 fine level of the multigrid V-cycle: the prolongation added into an
// FP32 xf (accumulated in FP64), and the first sweep from zero x = omega dinv b
// written in FP32
void launch_mg_prolong_add_f32(const Lattice &Lf, const Lattice &Lc, const double *xc,
                               const uint8_t *fmask_f, float *xf, hipStream_t s);
void launch_mg_jacobi0_f32(long long n, double omega, const double *dinv, const double *b,
                           float *x, hipStream_t s);
// coarse element values from the 8 children: mode 0 arithmetic mean, 1
// geometric mean, 3 quadratic mean, 4 maximum
void launch_mg_young(const Lattice &Lf, const Lattice &Lc, const double *yf, double *yc,
                     hipStream_t s, int mode = 0);
// damped Jacobi: y == nullptr: x = omega dinv b; else x += omega dinv (b - y)
// (nc load cases: b, y, x stacked with stride n, dinv shared)
void launch_mg_jacobi(long long n, double omega, const double *dinv, const double *b,
                      const double *y, double *x, hipStream_t s, int nc = 1);
// xout = xin + omega dinv (b - y) (xout may be xin) and the partials of
// xout.b and xout.q (the flexible PCG's z.r and z.q when this is a V-cycle's
// last sweep on r); returns their count
unsigned launch_mg_jacobi_dot(long long n, double omega, const double *dinv, const double *b,
                              const double *y, const double *xin, double *xout, const double *q,
                              double *part, hipStream_t s);

// One thread per node, K.x gathered from global memory (for coarse lattices,
// where a z-marching launch is latency-bound), with the consumer fused in:
enum MgDirectMode { MG_JACOBI = 1, MG_RESIDUAL = 2 };
//   MG_JACOBI:   out = x + omega dinv (b - K x)   (out may not alias x)
//   MG_RESIDUAL: out = b - K x
// every output masked by the free DOFs
void launch_mg_direct(int mode, const Lattice &L, const KeCoef &kc, const double *young,
                      const uint8_t *fmask, const double *x, const double *b, const double *dinv,
                      double omega, double *out, hipStream_t s, int nc = 1, long long cs = 0);
// The start of the W-cycle's second visit of a direct level in one pass:
// u = xc, r = b - K xc (masked) and, with x0 != nullptr, x0 = omega dinv r
// (xc may not alias u, r or x0)
void launch_mg_direct_wres(const Lattice &L, const KeCoef &kc, const double *young,
                           const uint8_t *fmask, const double *xc, const double *b,
                           const double *dinv, double omega, double *u, double *r, double *x0,
                           hipStream_t s, int nc = 1, long long cs = 0);

// out = xp + omega dinv (b - K xp) with xp = x + P xcoarse at the fine free
// DOFs (x elsewhere): launch_mg_prolong_add followed by an MG_JACOBI sweep, in
// one pass and without storing xp (out may not alias x).  With xcoarse2 the
// coarse values are xcoarse + xcoarse2 (the W-cycle's two coarse visits).
void launch_mg_prolong_jacobi(const Lattice &Lf, const Lattice &Lc, const KeCoef &kc,
                              const double *young, const uint8_t *fmask, const double *xcoarse,
                              const double *x, const double *b, const double *dinv, double omega,
                              double *out, hipStream_t s, const double *xcoarse2 = nullptr,
                              int nc = 1, long long sf = 0, long long sc = 0);

// Exact solve on a small lattice through the dense inverse of its K.  Dense
// index d = c * numnp + n for lattice DOF c * Np + n; constrained DOFs carry
// identity rows and columns.
constexpr int MG_DENSE_MAX = 6144;  // largest 3 numnp solved densely
// A = K of the lattice (n = 3 numnp, n x n, symmetric)
void launch_mg_dense_assemble(const Lattice &L, const double *ke, const double *young,
                              const uint8_t *fmask, double *A, hipStream_t s,
                              const double *kmat = nullptr);
// the diagonal of the free DOFs of the assembled dense A (n = 3 numnp) times (1 + shift)
void launch_mg_dense_shift(int n, const Lattice &L, const uint8_t *fmask, double shift, double *A,
                           hipStream_t s);
// A dense store holds the n x n matrix (column-major) and after it the packed
// tiles of its lower triangle: mg_dense_store_size(n) doubles.
long long mg_dense_store_size(long long n);
// packs the lower triangle of the store's matrix into its tiles
void launch_mg_dense_pack(int n, double *A, hipStream_t s);
// x = A b on the lattice (A symmetric: the store, read through its packed
// tiles), with a workspace part of mg_dense_part_size(L) doubles per load case.
// Launchers with nc > 1 run nc load cases in one launch: case c's vectors
// start at c x the level's stride (sf fine, sc coarse, cs both).
void launch_mg_dense_solve(const Lattice &L, const double *A, const double *b, double *x,
                           double *part, hipStream_t s, int nc = 1, long long cs = 0);
long long mg_dense_part_size(const Lattice &L);

}  // namespace icarat_hip
