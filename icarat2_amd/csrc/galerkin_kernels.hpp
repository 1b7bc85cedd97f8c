// galerkin_kernels.hpp -- launchers of the Galerkin coarse-operator kernels
// (k_galerkin.hip).  Element matrices are stored row-major, 576 doubles per
// element, elements in lattice order.
#pragma once
#include <hip/hip_runtime.h>

#include "device_common.hpp"
#include "hip_util.hpp"

namespace icarat_hip {

enum GalMode { GAL_APPLY = 0, GAL_RESIDUAL = 1, GAL_JACOBI = 2 };

// kc = the coarse level's element matrices R A_f P element by element, from the
// fine level's: young != nullptr: fine matrices young[e] * kf (kf = KE0);
// young == nullptr: explicit fine matrices kf.  Both levels' free masks apply.
// mu[C] = lambda_max(diag(K_C)^-1 K_C) over C's free DOFs (power iteration).
void launch_gal_build(const Lattice &Lf, const Lattice &Lc, const double *kf, const double *young,
                      const uint8_t *fmask_f, const uint8_t *fmask_c, double *kc, double *mu,
                      hipStream_t s);
// mu_max = max mu; dinv = theta / (mu_max diag(A)) at free DOFs, 0 elsewhere
void launch_gal_dinv(const Lattice &L, const double *km, const uint8_t *fmask, const double *mu,
                     double *mu_max, double theta, double *dinv, hipStream_t s);
// GAL_APPLY out = A x; GAL_RESIDUAL out = b - A x; GAL_JACOBI out = x + dinv (b - A x);
// masked by the free DOFs (out may not alias x)
void launch_gal_apply(int mode, const Lattice &L, const double *km, const uint8_t *fmask,
                      const double *x, const double *b, const double *dinv, double *out,
                      hipStream_t s);
// adds every element matrix into the dense 3 numnp x 3 numnp matrix A (row-major)
void launch_gal_dense_scatter(const Lattice &L, const double *km, double *A, hipStream_t s);

}  // namespace icarat_hip
