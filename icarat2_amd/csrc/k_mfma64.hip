// k_mfma64.hip -- FP64 K.u on the matrix cores: element-centric contraction
// F_e = KE0 (E_e u_e) with v_mfma_f64_4x4x4_4b (four independent 4x4x4 FP64
// products per instruction), then a node-centric gather of the element forces.
//
// KE0 (24x24) is split into 6x6 blocks of 4x4.  For a group of 16 elements
// (lanes l: element l % 16, row/column quarter q = l / 16), one instruction
// applies block (ib, kb) to four elements per 4-lane block, so 36 instructions
// give the 24 force components of 16 elements: 576 FMAs per element, no
// padding.  Lane l holds, for every block,
//   A: KE0[4 ib + l % 4][4 kb + q]     (the same for the four blocks: registers)
//   B: E_e u_e[4 kb + q]               (element l % 16, from the LDS node planes)
//   D: F_e[4 ib + q]                   (element l % 16)
// (the operand layout of the instruction as measured on gfx950: block (l % 16) / 4,
// A row l % 4 and column l / 16, B row l / 16 and column l % 4, D row l / 16).
//
// The tile is the z-marching K.u's: a workgroup owns a 32 x 8 node column and
// marches over the element layers of one z-chunk.  Per layer the 33 x 9
// elements around the column (one halo ring in x and y, computed redundantly
// with the neighbouring tiles) are contracted from the two node planes in LDS;
// their forces go to LDS, and each thread sums, in a fixed element order, the
// contributions to its bottom node (finished, written) and its top node (kept
// for the next layer).  Deterministic, each node written once.
#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace icarat_hip {

namespace {

constexpr int NT = 256;
constexpr int ZX = 32, ZY = 8;
constexpr int ZPX = ZX + 2, ZPY = ZY + 2, ZPN = ZPX * ZPY;  // halo node plane
constexpr int ZPER = (ZPN + NT - 1) / NT;
constexpr int EXL = ZX + 1, EYL = ZY + 1, NEL = EXL * EYL;   // elements per layer
constexpr int NGRP = (NEL + 15) / 16;                        // groups of 16 elements
constexpr int NELP = NGRP * 16;
constexpr int EPER = (NELP + NT - 1) / NT;                   // layer elements per thread

__device__ __forceinline__ unsigned xcd_tile(unsigned b, unsigned nt) {
  const unsigned q = nt / 8, rem = nt % 8, xcd = b % 8, slot = b / 8;
  return (xcd < rem ? xcd * (q + 1) : rem * (q + 1) + (xcd - rem) * q) + slot;
}

__global__ __launch_bounds__(NT, 2) void k_apply64_mfma(Lattice L, int nchunks,
                                                       const double *__restrict__ KE0,
                                                       const double *__restrict__ young,
                                                       const uint8_t *__restrict__ fmask,
                                                       const double *__restrict__ src,
                                                       double *__restrict__ y) {
  __shared__ double sp[2][3][ZPN];
  __shared__ double se[2][NELP];
  __shared__ double sF[24][NELP];
  const int nbx = (L.NX + ZX - 1) / ZX, nby = (L.NY + ZY - 1) / ZY, ncols = nbx * nby;
  const unsigned T_ = xcd_tile(blockIdx.x, gridDim.x);
  const int col = (int)(T_ % ncols), zc = (int)(T_ / ncols);
  const int x0 = (col % nbx) * ZX, y0 = (col / nbx) * ZY;
  const int z0 = (int)((long long)zc * L.NZ / nchunks);
  const int z1 = (int)((long long)(zc + 1) * L.NZ / nchunks);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int e16 = lane & 15, q = lane >> 4;
  // A operands: KE0 block (ib, kb), row 4 ib + lane % 4, column 4 kb + q
  double A[6][6];
#pragma unroll
  for (int ib = 0; ib < 6; ++ib)
#pragma unroll
    for (int kb = 0; kb < 6; ++kb) A[ib][kb] = KE0[(4 * ib + (lane & 3)) * 24 + 4 * kb + q];
  // B operand k = 4 kb + q: local node k / 3 (its plane offset from the
  // element's base node and its z side) and component k % 3
  int boff[6], bz[6];
#pragma unroll
  for (int kb = 0; kb < 6; ++kb) {
    const int k = 4 * kb + q, ln = k / 3, c = k % 3;
    boff[kb] = c * ZPN + lnode_x(ln) + ZPX * lnode_y(ln);
    bz[kb] = lnode_z(ln);
  }
  const int tx = threadIdx.x % ZX, ty = threadIdx.x / ZX;
  const int gx = x0 + tx, gy = y0 + ty;
  const bool col_in = gx < L.NX && gy < L.NY;

  double pf[ZPER][3], pe[EPER];
  // node plane zp and element layer zp - 1 (the layer whose top plane is zp)
  auto fetch = [&](int zp) {
#pragma unroll
    for (int k = 0; k < ZPER; ++k) {
      const int t = threadIdx.x + k * NT;
      const int hx = x0 - 1 + t % ZPX, hy = y0 - 1 + t / ZPX;
      const bool in = t < ZPN && zp >= 0 && zp < L.NZ && hx >= 0 && hx < L.NX && hy >= 0 &&
                      hy < L.NY;
      const long long n = in ? hx + (long long)L.NX * (hy + (long long)L.NY * zp) : 0;
#pragma unroll
      for (int c = 0; c < 3; ++c) pf[k][c] = in ? src[c * L.Np + n] : 0.0;
    }
    const int ze = zp - 1;
#pragma unroll
    for (int k = 0; k < EPER; ++k) {
      const int t = threadIdx.x + k * NT;
      const int ex = x0 - 1 + t % EXL, ey = y0 - 1 + t / EXL;
      const bool in = t < NEL && ze >= 0 && ze < L.nz && ex >= 0 && ex < L.nx && ey >= 0 &&
                      ey < L.ny;
      pe[k] = in ? young[ex + (long long)L.nx * (ey + (long long)L.ny * ze)] : 0.0;
    }
  };
  // planes go to slot s, the element layer to slot s ^ 1 (it is used with its
  // bottom plane's slot)
  auto commit = [&](int slot) {
#pragma unroll
    for (int k = 0; k < ZPER; ++k) {
      const int t = threadIdx.x + k * NT;
      if (t < ZPN) {
#pragma unroll
        for (int c = 0; c < 3; ++c) sp[slot][c][t] = pf[k][c];
      }
    }
#pragma unroll
    for (int k = 0; k < EPER; ++k) {
      const int t = threadIdx.x + k * NT;
      if (t < NELP) se[slot ^ 1][t] = pe[k];
    }
  };
  // planes z0-1 -> slot 0, z0 -> slot 1 (layer step k reads slots k & 1 and its other)
  fetch(z0 - 1);
  commit(0);
  fetch(z0);
  commit(1);
  __syncthreads();
  double acc[3] = {0.0, 0.0, 0.0};
  const int nsteps = z1 - z0 + 1;  // element layers z0-1 .. z1-1
  for (int k = 0; k < nsteps; ++k) {
    const int zb = z0 - 1 + k;
    const int s0 = k & 1, s1 = s0 ^ 1;
    const bool more = k + 1 < nsteps;
    if (more) fetch(zb + 2);
    // element forces of layer zb (young in se[s0]: zero outside the mesh and
    // for the padding of the last group)
    for (int g = wave; g < NGRP; g += NT / 64) {
      const int el = 16 * g + e16;
      const int base = el < NEL ? el % EXL + ZPX * (el / EXL) : 0;
      const double E = se[s0][el];
      double B[6];
#pragma unroll
      for (int kb = 0; kb < 6; ++kb) {
        const double *pl = bz[kb] ? &sp[s1][0][0] : &sp[s0][0][0];
        B[kb] = E * pl[base + boff[kb]];
      }
      // six independent accumulators, one per row block
      double d[6];
#pragma unroll
      for (int ib = 0; ib < 6; ++ib) d[ib] = 0.0;
#pragma unroll
      for (int kb = 0; kb < 6; ++kb)
#pragma unroll
        for (int ib = 0; ib < 6; ++ib)
          d[ib] = __builtin_amdgcn_mfma_f64_4x4x4f64(A[ib][kb], B[kb], d[ib], 0, 0, 0);
#pragma unroll
      for (int ib = 0; ib < 6; ++ib) sF[4 * ib + q][el] = d[ib];
    }
    __syncthreads();
    // gather: the layer's 4 elements around the column, fixed order
    double bot[3] = {0.0, 0.0, 0.0}, top[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      const int ox = s & 1, oy = s >> 1;           // element (tx + ox, ty + oy) of the layer
      const int el = (tx + ox) + EXL * (ty + oy);
      const int lx = 1 - ox, ly = 1 - oy;           // the node's corner in that element
      const int lb = lnode(lx, ly, 0), lt = lnode(lx, ly, 1);
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        bot[c] += sF[3 * lb + c][el];
        top[c] += sF[3 * lt + c][el];
      }
    }
    if (k >= 1 && col_in) {
      const long long n = gx + (long long)L.NX * (gy + (long long)L.NY * zb);
      const unsigned m = fmask[n];
#pragma unroll
      for (int c = 0; c < 3; ++c) y[c * L.Np + n] = ((m >> c) & 1u) ? acc[c] + bot[c] : 0.0;
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) acc[c] = top[c];
    if (more) commit(s0);
    __syncthreads();
  }
}

}  // namespace

void launch_apply_k64_mfma(const Lattice &L, const KeCoef &ke, const double *young,
                           const uint8_t *freemask, const double *x, double *y, hipStream_t s) {
  const int def = stencil_nchunks(L);
  const int nch = ke.nchunks > 0 ? std::min(ke.nchunks, def) : def;
  const int ncols = ((L.NX + ZX - 1) / ZX) * ((L.NY + ZY - 1) / ZY);
  hipLaunchKernelGGL(k_apply64_mfma, dim3((unsigned)(ncols * nch)), dim3(NT), 0, s, L, nch, ke.ke,
                     young, freemask, x, y);
  HIPCHK(hipGetLastError());
}

}  // namespace icarat_hip
