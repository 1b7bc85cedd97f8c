// contract.hpp -- the node-centric hexa8 contraction shared by the K.u
// kernels: for one lattice node, the 3 rows of K.u gathered from the 27
// neighbouring nodal displacements and the 8 surrounding elements' young.
#pragma once
#include <hip/hip_runtime.h>

#include "device_common.hpp"
#include "kernels.hpp"
#include "ke_sym_table.hpp"

namespace icarat_hip {

__device__ __forceinline__ constexpr int nb(int dx, int dy, int dz) {
  return (dx + 1) + 3 * (dy + 1) + 9 * (dz + 1);
}

// KE0[e] (e = row*24 + col, a compile-time constant after unrolling): from
// memory (SYM 0) or as +-orbit value held in scalar registers (SYM 1 box, 2 cube).
template <int SYM, class T>
__device__ __forceinline__ T kcoef(const T *__restrict__ c_KE0, const KeValsT<T> &kv, int e) {
  if constexpr (SYM == 0) {
    return c_KE0[e];
  } else {
    const int code = SYM == 1 ? KSYM_BOX[e] : KSYM_CUBE[e];
    return code > 0 ? kv.v[code - 1] : -kv.v[-code - 1];
  }
}

// y = sum over the 8 elements around a node of E_s * KE0[rows of the node] . u_e
template <int SYM>
__device__ __forceinline__ void node_contract(const double *__restrict__ c_KE0, const KeVals &kv,
                                              const double (&U)[27][3], const double (&E)[8],
                                              double (&y)[3]) {
  y[0] = y[1] = y[2] = 0.0;
#pragma unroll
  for (int s = 0; s < 8; ++s) {
    const int ex = (s & 1) - 1, ey = ((s >> 1) & 1) - 1, ez = ((s >> 2) & 1) - 1;
    const int a = lnode(-ex, -ey, -ez);
    double t0 = 0.0, t1 = 0.0, t2 = 0.0;
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const int o = nb(ex + lnode_x(b), ey + lnode_y(b), ez + lnode_z(b));
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        const double u = U[o][j];
        t0 += kcoef<SYM>(c_KE0, kv, (3 * a + 0) * 24 + 3 * b + j) * u;
        t1 += kcoef<SYM>(c_KE0, kv, (3 * a + 1) * 24 + 3 * b + j) * u;
        t2 += kcoef<SYM>(c_KE0, kv, (3 * a + 2) * 24 + 3 * b + j) * u;
      }
    }
    y[0] += E[s] * t0;
    y[1] += E[s] * t1;
    y[2] += E[s] * t2;
  }
}

}  // namespace icarat_hip
