// mixed_kernels.hpp -- launchers of the MFMA contraction and the FP32 inner
// PCG / FP64 refinement kernels (k_mixed.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "device_common.hpp"
#include "hip_util.hpp"

namespace icarat_hip {

// mode 0: FP32 operands on v_mfma_f32_16x16x4_f32; mode 1: BF16 operands on
// v_mfma_f32_16x16x32_bf16.  Accumulation in FP32 in both.
int mixed_zchunk(const Lattice &L);
long long mixed_blocks(const Lattice &L);
void launch_apply_k_mixed(const Lattice &L, int mode, const float *ke, const float *young,
                          const uint8_t *fmask, const float *u, float *y, hipStream_t s);
void launch_pcg_stencil_mixed(const Lattice &L, int mode, const float *ke, const float *young,
                              const uint8_t *fmask, const float *r, const float *dinv,
                              const float *p_old, float *p_new, float *q, PcgScalars *S,
                              double *partials, hipStream_t s);
void launch_pcg_update32(long long n, float *d, float *r, const float *p, const float *q,
                         const float *dinv, PcgScalars *S, double *partials, hipStream_t s);
// r32 = b - y (y = K64 x64 or null); init: also d = p0 = 0 and the scalars reset
void launch_replace32(long long n, const double *b, const double *y, float *r, float *d,
                      float *p0, const float *dinv, PcgScalars *S, double *partials, int init,
                      hipStream_t s);
// x64 += d32; d32 = 0
void launch_ir_accumulate(long long n, float *d, double *x, hipStream_t s);
void launch_f64_to_f32(long long n, const double *a, float *b, hipStream_t s);
void launch_f32_to_f64(long long n, const float *a, double *b, hipStream_t s);

}  // namespace icarat_hip
