// mg_kernels.hpp -- launchers of the multigrid kernels (k_mg.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "device_common.hpp"
#include "hip_util.hpp"

namespace icarat_hip {

// bc = R rf on the coarse lattice (transpose of trilinear prolongation), masked
void launch_mg_restrict(const Lattice &Lf, const Lattice &Lc, const double *rf,
                        const uint8_t *fmask_c, double *bc, hipStream_t s);
// xf += P xc (trilinear interpolation), masked
void launch_mg_prolong_add(const Lattice &Lf, const Lattice &Lc, const double *xc,
                           const uint8_t *fmask_f, double *xf, hipStream_t s);
// coarse Young's moduli: mean over the 8 children
void launch_mg_young(const Lattice &Lf, const Lattice &Lc, const double *yf, double *yc,
                     hipStream_t s);
// damped Jacobi: y == nullptr: x = omega dinv b; else x += omega dinv (b - y)
void launch_mg_jacobi(long long n, double omega, const double *dinv, const double *b,
                      const double *y, double *x, hipStream_t s);
void launch_mg_residual(long long n, const double *b, const double *y, double *r, hipStream_t s);

}  // namespace icarat_hip
