// k_mixed.hip -- K·u with the element contraction on the matrix cores (MFMA)
// in FP32 or BF16 with FP32 accumulation: the matvec of the mixed-precision
// Jacobi-PCG, whose vectors, scalars and reductions stay FP64 and whose FP64
// true residual periodically replaces the recursive one.
//
// The vectors stay FP64 (the kernel reads FP64 r, dinv, p_old and writes FP64
// p and q); only the staged node planes and the contraction are FP32 / BF16.
//
// Element-centric contraction: W = KE0 · U, with KE0 (24x24, rows padded to
// 32) as the MFMA A operand held in registers for the whole kernel and U the
// gathered element displacement vectors of 16 elements per MFMA tile (the N
// dimension).  A workgroup owns a 32x8 column of nodes and marches over the
// element layers in z: for each layer it builds the B operands from an LDS
// ring of node planes, contracts on MFMA, writes the young-scaled element
// forces to LDS and then every node sums the forces of its 4 elements in the
// layer (no atomics; deterministic).
//
// U enters as differences to the element's local node 0 (u_b - u_0 per
// component).  KE0 annihilates rigid translations, so K·u is unchanged in
// exact arithmetic, while the rounding of U to FP32/BF16 now scales with the
// displacement differences across one element instead of the displacements.
//
// Operand layouts (one 64-lane wave; verified on gfx950 by a lane probe):
//   v_mfma_f32_16x16x4_f32 : A[i][k] lane i+16k, B[k][j] lane j+16k,
//                            D[4*(l/16)+v][l%16] in accumulator v
//   v_mfma_f32_16x16x32_bf16: A[i][8*(l/16)+v] and B[8*(l/16)+v][j] in element v
//                            of lane l (i, j = l%16), D as above
#include <hip/hip_runtime.h>

#include "mixed_kernels.hpp"

namespace icarat_hip {

namespace {

typedef float v4f __attribute__((ext_vector_type(4)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));

constexpr int TX = 32, TY = 8, NT = 256;
constexpr int PX = TX + 2, PY = TY + 2, PN = PX * PY;   // node plane window
constexpr int LX = TX + 1, LY = TY + 1, LN = LX * LY;   // elements per layer
constexpr int NTILE = (LN + 15) / 16;                   // 16-element MFMA tiles
constexpr int WSTRIDE = 25;                             // padded row of sW
constexpr int VEC_BLOCKS = 2048;

// local node b of an element: offsets in the plane window / bottom-top
__device__ __forceinline__ int bx_(int b) { return lnode_x(b); }
__device__ __forceinline__ int by_(int b) { return lnode_y(b); }
__device__ __forceinline__ int bz_(int b) { return lnode_z(b); }

template <int MODE>
struct KeOperand;

// FP32: 2 row tiles x 6 k-steps of 4, one float per lane each
template <>
struct KeOperand<0> {
  float a[2][6];
  __device__ void load(const float *__restrict__ ke) {
    const int l = threadIdx.x & 63;
#pragma unroll
    for (int m = 0; m < 2; ++m)
#pragma unroll
      for (int s = 0; s < 6; ++s) {
        const int i = 16 * m + (l & 15), k = 4 * s + (l >> 4);
        a[m][s] = i < 24 ? ke[i * 24 + k] : 0.0f;
      }
  }
};

// BF16: 2 row tiles x 1 k-step of 32 (24 used), 8 bf16 per lane each
template <>
struct KeOperand<1> {
  v8bf a[2];
  __device__ void load(const float *__restrict__ ke) {
    const int l = threadIdx.x & 63;
#pragma unroll
    for (int m = 0; m < 2; ++m)
#pragma unroll
      for (int v = 0; v < 8; ++v) {
        const int i = 16 * m + (l & 15), k = 8 * (l >> 4) + v;
        a[m][v] = (__bf16)((i < 24 && k < 24) ? ke[i * 24 + k] : 0.0f);
      }
  }
};

// value of DOF d (= 3*b + c) of layer element (lx,ly) minus that of its node 0
__device__ __forceinline__ float elem_dof(const float *pl0, const float *pl1, int lx, int ly, int d) {
  const int b = d / 3, c = d % 3;
  const float *pb = bz_(b) ? pl1 : pl0;
  const float v = pb[c * PN + (lx + bx_(b)) + PX * (ly + by_(b))];
  const float v0 = pl0[c * PN + lx + PX * ly];
  return v - v0;
}

// Contract all element tiles of one layer; write young-scaled forces to sW.
template <int MODE>
__device__ __forceinline__ void contract_layer(const KeOperand<MODE> &K, const float *pl0,
                                               const float *pl1, const float *sE, float *sW) {
  const int l = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (int t = w; t < NTILE; t += NT / 64) {
    const int le = 16 * t + (l & 15);
    const bool valid = le < LN;
    const int lx = valid ? le % LX : 0, ly = valid ? le / LX : 0;
    v4f acc[2] = {{0.f, 0.f, 0.f, 0.f}, {0.f, 0.f, 0.f, 0.f}};
    if constexpr (MODE == 0) {
#pragma unroll
      for (int s = 0; s < 6; ++s) {
        const int d = 4 * s + (l >> 4);
        const float bval = valid ? elem_dof(pl0, pl1, lx, ly, d) : 0.0f;
        acc[0] = __builtin_amdgcn_mfma_f32_16x16x4f32(K.a[0][s], bval, acc[0], 0, 0, 0);
        acc[1] = __builtin_amdgcn_mfma_f32_16x16x4f32(K.a[1][s], bval, acc[1], 0, 0, 0);
      }
    } else {
      v8bf bv;
#pragma unroll
      for (int v = 0; v < 8; ++v) {
        const int d = 8 * (l >> 4) + v;
        bv[v] = (__bf16)((valid && d < 24) ? elem_dof(pl0, pl1, lx, ly, d) : 0.0f);
      }
      acc[0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(K.a[0], bv, acc[0], 0, 0, 0);
      acc[1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(K.a[1], bv, acc[1], 0, 0, 0);
    }
    if (valid) {
      const float E = sE[le];
#pragma unroll
      for (int m = 0; m < 2; ++m)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int i = 16 * m + 4 * (l >> 4) + v;
          if (i < 24) sW[le * WSTRIDE + i] = E * acc[m][v];
        }
    }
  }
}

// Loads node plane z of the window into LDS (3 components, zero outside the
// mesh).  PCG: the plane is the new search direction p = dinv*r + beta*p_old,
// written back for owned nodes.
template <bool PCG, class VT>
__device__ __forceinline__ void load_plane(const Lattice &L, int x0, int y0, int z, bool own_z,
                                           const VT *__restrict__ src,
                                           const VT *__restrict__ dinv,
                                           const VT *__restrict__ p_old,
                                           VT *__restrict__ p_new, VT beta, float *pl) {
  for (int t = threadIdx.x; t < PN; t += NT) {
    const int hx = t % PX, hy = t / PX;
    const int gx = x0 - 1 + hx, gy = y0 - 1 + hy;
    VT v[3] = {VT(0), VT(0), VT(0)};
    if (z >= 0 && z < L.NZ && gx >= 0 && gx < L.NX && gy >= 0 && gy < L.NY) {
      const long long n = gx + (long long)L.NX * (gy + (long long)L.NY * z);
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const long long q = c * L.Np + n;
        if constexpr (PCG) v[c] = dinv[q] * src[q] + beta * p_old[q];
        else v[c] = src[q];
      }
      if constexpr (PCG) {
        if (own_z && hx >= 1 && hx <= TX && hy >= 1 && hy <= TY)
#pragma unroll
          for (int c = 0; c < 3; ++c) p_new[c * L.Np + n] = v[c];
      }
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) pl[c * PN + t] = (float)v[c];
  }
}

template <int MODE, bool PCG, class VT>
__global__ __launch_bounds__(NT) void k_stencil_mfma(
    Lattice L, int zchunk, const float *__restrict__ ke, const float *__restrict__ young,
    const uint8_t *__restrict__ fmask, const VT *__restrict__ src,
    const VT *__restrict__ dinv, const VT *__restrict__ p_old, VT *__restrict__ p_new,
    VT *__restrict__ y, PcgScalars *S, double *partials, unsigned *counter) {
  __shared__ float sP[3][3 * PN];       // ring of node planes
  __shared__ float sW[LN * WSTRIDE];    // element forces of the current layer
  __shared__ float sE[LN];              // young of the current layer
  __shared__ double sh[4];
  VT beta = VT(0);
  if constexpr (PCG) {
    if (S->done) return;
    beta = (VT)S->beta;
  }
  const int x0 = blockIdx.x * TX, y0 = blockIdx.y * TY, z0 = blockIdx.z * zchunk;
  const int z1 = min(z0 + zchunk, L.NZ);  // owned node planes [z0, z1)
  KeOperand<MODE> K;
  K.load(ke);
  const int tx = threadIdx.x % TX, ty = threadIdx.x / TX;
  const int gx = x0 + tx, gy = y0 + ty;
  const bool node_in = gx < L.NX && gy < L.NY;
  const long long nxy = gx + (long long)L.NX * gy;
  float acc[3] = {0.f, 0.f, 0.f};
  double pq = 0.0;
  // planes z0-1 and z0 into the ring
  load_plane<PCG, VT>(L, x0, y0, z0 - 1, false, src, dinv, p_old, p_new, beta, sP[0]);
  load_plane<PCG, VT>(L, x0, y0, z0, z0 < z1, src, dinv, p_old, p_new, beta, sP[1]);
  for (int lay = 0; lay <= z1 - z0; ++lay) {
    const int zb = z0 - 1 + lay;           // element layer between planes zb, zb+1
    const float *pl0 = sP[lay % 3], *pl1 = sP[(lay + 1) % 3];
    // young of this layer (0 outside the mesh)
    for (int t = threadIdx.x; t < LN; t += NT) {
      const int ex = x0 - 1 + t % LX, ey = y0 - 1 + t / LX;
      float v = 0.f;
      if (zb >= 0 && zb < L.nz && ex >= 0 && ex < L.nx && ey >= 0 && ey < L.ny)
        v = young[ex + (long long)L.nx * (ey + (long long)L.ny * zb)];
      sE[t] = v;
    }
    __syncthreads();
    contract_layer<MODE>(K, pl0, pl1, sE, sW);
    // prefetch plane zb+2 for the next layer (the ring slot of plane zb-1 is free)
    if (lay + 1 <= z1 - z0)
      load_plane<PCG, VT>(L, x0, y0, zb + 2, zb + 2 < z1, src, dinv, p_old, p_new, beta,
                          sP[(lay + 2) % 3]);
    __syncthreads();
    // gather: plane zb is complete (acc from the layer below + bottom forces here)
    float bot[3] = {0.f, 0.f, 0.f}, top[3] = {0.f, 0.f, 0.f};
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int ex = (e & 1) - 1, ey = (e >> 1) - 1;
      const int le = (tx + 1 + ex) + LX * (ty + 1 + ey);
      const int ab = lnode(-ex, -ey, 0), at = lnode(-ex, -ey, 1);
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        bot[c] += sW[le * WSTRIDE + 3 * ab + c];
        top[c] += sW[le * WSTRIDE + 3 * at + c];
      }
    }
    if (lay >= 1 && node_in && zb < L.NZ) {
      const long long n = nxy + (long long)L.NX * L.NY * zb;
      const unsigned mk = fmask[n];
      float out[3];
#pragma unroll
      for (int c = 0; c < 3; ++c) out[c] = ((mk >> c) & 1u) ? acc[c] + bot[c] : 0.f;
#pragma unroll
      for (int c = 0; c < 3; ++c) y[c * L.Np + n] = (VT)out[c];
      if constexpr (PCG) {
        const int h = (tx + 1) + PX * (ty + 1);
#pragma unroll
        for (int c = 0; c < 3; ++c) pq += (double)pl0[c * PN + h] * (double)out[c];
      }
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) acc[c] = top[c];
    __syncthreads();
  }
  if constexpr (PCG) {
    double v[1] = {pq};
    store_partials<NT, 1>(v, partials, sh);
  }
}

__global__ void k_f64_to_f32(long long n, const double *__restrict__ a, float *__restrict__ b) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i < n) b[i] = (float)a[i];
}
__global__ void k_f32_to_f64(long long n, const float *__restrict__ a, double *__restrict__ b) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i < n) b[i] = (double)a[i];
}

inline unsigned blocks_for(long long n, int nt) { return (unsigned)((n + nt - 1) / nt); }

}  // namespace

int mixed_zchunk(const Lattice &L) {
  // enough workgroups to fill 256 CUs several times over
  const long long cols = (long long)((L.NX + TX - 1) / TX) * ((L.NY + TY - 1) / TY);
  int zc = 16;
  while (zc > 4 && cols * ((L.NZ + zc - 1) / zc) < 2048) zc /= 2;
  return zc;
}

static dim3 mixed_grid(const Lattice &L, int zc) {
  return dim3((L.NX + TX - 1) / TX, (L.NY + TY - 1) / TY, (L.NZ + zc - 1) / zc);
}

long long mixed_blocks(const Lattice &L) {
  const int zc = mixed_zchunk(L);
  dim3 g = mixed_grid(L, zc);
  return (long long)g.x * g.y * g.z;
}

void launch_apply_k_mixed(const Lattice &L, int mode, const float *ke, const float *young,
                          const uint8_t *fmask, const double *u, double *y, hipStream_t s) {
  const int zc = mixed_zchunk(L);
  if (mode == 0)
    hipLaunchKernelGGL((k_stencil_mfma<0, false, double>), mixed_grid(L, zc), dim3(NT), 0, s, L, zc,
                       ke, young, fmask, u, nullptr, nullptr, nullptr, y, nullptr, nullptr, nullptr);
  else
    hipLaunchKernelGGL((k_stencil_mfma<1, false, double>), mixed_grid(L, zc), dim3(NT), 0, s, L, zc,
                       ke, young, fmask, u, nullptr, nullptr, nullptr, y, nullptr, nullptr, nullptr);
  HIPCHK(hipGetLastError());
}

void launch_pcg_stencil_mfma64(const Lattice &L, int mode, const float *ke, const float *young,
                               const uint8_t *fmask, const double *r, const double *dinv,
                               const double *p_old, double *p_new, double *q, PcgScalars *S,
                               double *partials, hipStream_t s, bool finalize) {
  const int zc = mixed_zchunk(L);
  const dim3 g = mixed_grid(L, zc);
  if (mode == 0)
    hipLaunchKernelGGL((k_stencil_mfma<0, true, double>), g, dim3(NT), 0, s, L, zc, ke, young, fmask,
                       r, dinv, p_old, p_new, q, S, partials, nullptr);
  else
    hipLaunchKernelGGL((k_stencil_mfma<1, true, double>), g, dim3(NT), 0, s, L, zc, ke, young, fmask,
                       r, dinv, p_old, p_new, q, S, partials, nullptr);
  HIPCHK(hipGetLastError());
  if (finalize) launch_finalize(kFinPcgAlpha, 1, partials, g.x * g.y * g.z, S, nullptr, nullptr, s);
}

void launch_f64_to_f32(long long n, const double *a, float *b, hipStream_t s) {
  hipLaunchKernelGGL(k_f64_to_f32, dim3(blocks_for(n, 256)), dim3(256), 0, s, n, a, b);
  HIPCHK(hipGetLastError());
}

void launch_f32_to_f64(long long n, const float *a, double *b, hipStream_t s) {
  hipLaunchKernelGGL(k_f32_to_f64, dim3(blocks_for(n, 256)), dim3(256), 0, s, n, a, b);
  HIPCHK(hipGetLastError());
}

}  // namespace icarat_hip
