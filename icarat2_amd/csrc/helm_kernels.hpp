// helm_kernels.hpp -- launchers of the Helmholtz filter kernels (k_helm.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "device_common.hpp"
#include "hip_util.hpp"

namespace icarat_hip {

// rhs[a] = s * sum of in[e] over the elements that share node a (numnp)
void launch_helm_rhs(const Lattice &L, const double *in, double s, double *rhs, hipStream_t st);
// out[e] = s * sum of val over the 8 nodes of element e (nelx)
void launch_helm_extract(const Lattice &L, const double *val, double s, double *out,
                         hipStream_t st);
// v[i,j,k] /= 1 + c (lx[i] + ly[j] + lz[k]) on the node lattice
void launch_helm_scale(const Lattice &L, const double *lx, const double *ly, const double *lz,
                       double c, double *v, hipStream_t st);

}  // namespace icarat_hip
