// mixed_kernels.hpp -- launchers of the MFMA contraction kernels (k_mixed.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "device_common.hpp"
#include "hip_util.hpp"

namespace icarat_hip {

// mode 0: FP32 operands on v_mfma_f32_16x16x4_f32; mode 1: BF16 operands on
// v_mfma_f32_16x16x32_bf16.  Accumulation in FP32 in both.
int mixed_zchunk(const Lattice &L);
long long mixed_blocks(const Lattice &L);
// y = K u with FP64 vectors and the low-precision contraction
void launch_apply_k_mixed(const Lattice &L, int mode, const float *ke, const float *young,
                          const uint8_t *fmask, const double *u, double *y, hipStream_t s);
// FP64-vector PCG stencil half with the low-precision contraction (the
// matvec of the mixed-precision solve; vectors, scalars and reductions FP64)
void launch_pcg_stencil_mfma64(const Lattice &L, int mode, const float *ke, const float *young,
                               const uint8_t *fmask, const double *r, const double *dinv,
                               const double *p_old, double *p_new, double *q, PcgScalars *S,
                               double *partials, hipStream_t s, bool finalize = true);
void launch_f64_to_f32(long long n, const double *a, float *b, hipStream_t s);
void launch_f32_to_f64(long long n, const float *a, double *b, hipStream_t s);

}  // namespace icarat_hip
