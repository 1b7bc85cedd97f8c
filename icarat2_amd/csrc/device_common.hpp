// device_common.hpp -- shared device helpers: lattice geometry, the hexa8
// local-node table, deterministic wave/block reductions and the
// "last block finalises" pattern used by every reduction in the library.
#pragma once
#include <cmath>
#include <hip/hip_runtime.h>

#include <cstdint>

namespace icarat_hip {

// Node lattice of an nx*ny*nz voxel mesh.  Node n = i + NX*(j + NY*k) with
// NX = nx+1 (x fastest), element e = i + nx*(j + ny*k); DOF vectors are
// stored component-major: v[c*Np + n], Np = numnp padded to 256.
struct Lattice {
  int nx, ny, nz;   // elements
  int NX, NY, NZ;   // nodes
  long long numnp, nelx, Np;
};

// Local hexa8 node index of the element corner at offset (ox,oy,oz) in
// {0,1}^3 from the element's minimum corner: bottom face counter-clockwise
// 0..3, top face 4..7 (reference io/mesh_icarat.hpp:486-509).
__host__ __device__ constexpr int lnode(int ox, int oy, int oz) {
  return (ox == 0 ? (oy == 0 ? 0 : 3) : (oy == 0 ? 1 : 2)) + 4 * oz;
}
// x^p of the SIMP law (problem/elasticity.hpp:13-17).  For the usual p = 3
// the cube is formed with the exact error terms of its two products, which
// gives the correctly rounded x^3 (the value a correctly rounded pow returns)
// on the device and on the host alike; any other p goes to pow.
__host__ __device__ inline double simp_pow(double x, double p) {
  if (p == 3.0) {
    const double s = x * x;
    const double es = fma(x, x, -s);  // x^2 = s + es exactly
    const double t = s * x;
    const double et = fma(s, x, -t);  // s x = t + et exactly
    return t + fma(es, x, et);
  }
  return pow(x, p);
}

__host__ __device__ constexpr int lnode_x(int a) { return ((a & 3) == 1 || (a & 3) == 2) ? 1 : 0; }
__host__ __device__ constexpr int lnode_y(int a) { return ((a & 3) >= 2) ? 1 : 0; }
__host__ __device__ constexpr int lnode_z(int a) { return a >> 2; }

// Scalars of one device-resident Jacobi-PCG solve.
struct PcgScalars {
  double rho;        // r.z of the current iterate
  double pq;         // p.Ap
  double rr;         // r.r
  double bnorm2;     // b.b
  double tol;        // relative tolerance on ||r|| / ||b||
  double alpha, beta;
  double delta;      // reliable updates: replace r once ||r|| < delta * ||r at last replacement||
  double repl_norm;  // current replacement threshold (0: none)
  int iter, maxit;
  int done;          // 1 once converged, capped, broken down or a replacement is due
  int status;        // 0 running/converged, 1 capped, 2 breakdown, 3 replacement due,
                     // 4 preconditioned residual not a descent direction
  int restart;       // flexible PCG: next beta = 0
  // Fused Jacobi-PCG batches (the finalizes folded into the next kernel):
  // rho of in-batch iteration j lives in rho_slot[j & 1] so that a kernel never
  // reads a field its own blocks write; iter_base = iterations before the batch.
  int iter_base;
  double rho_slot[2];
};

// Scalars of the device-resident OC bisection.
struct OcState {
  double l0, l1, lam, lam_last, g, vol;
  int counter, done;
  int maxiter;
  int pad;
};

// --- deterministic reductions --------------------------------------------
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Sum over a block of NT threads; result valid on every thread.
template <int NT, int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV], double *sh /* NV*NT/64 */) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    double s = wave_sum(v[k]);
    if (lane == 0) sh[k * (NT / 64) + w] = s;
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < NT / 64; ++i) s += sh[k * (NT / 64) + i];
    v[k] = s;
  }
  __syncthreads();
}

// Every block stores its partial sums (plain stores, no fences); a separate
// single-block finalize kernel sums them in block order after the kernel
// boundary has made them visible.  Deterministic for a fixed grid.
// (store_partials_at: the block's partials go to slot bid of G per quantity)
template <int NT, int NV>
__device__ __forceinline__ void store_partials_at(double (&v)[NV], double *partials, double *sh,
                                                  unsigned G, unsigned bid) {
  block_sum<NT, NV>(v, sh);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) partials[(size_t)k * G + bid] = v[k];
  }
}

template <int NT, int NV>
__device__ __forceinline__ void store_partials(double (&v)[NV], double *partials, double *sh) {
  const unsigned G = gridDim.x * gridDim.y * gridDim.z;
  const unsigned bid = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
  store_partials_at<NT, NV>(v, partials, sh, G, bid);
}

// Fixed-order sum of G partials per quantity (call from one block of NT threads).
// Each thread issues up to 8 loads per quantity before adding them, so the
// latency of the partials (written by the previous kernel, possibly on another
// XCD) is paid once per group rather than once per value.
template <int NT, int NV>
__device__ __forceinline__ void sum_partials(const double *partials, unsigned G, double (&tot)[NV],
                                             double *sh) {
  constexpr int U = 8;
#pragma unroll
  for (int k = 0; k < NV; ++k) tot[k] = 0.0;
  for (unsigned base = 0; base < G; base += NT * U) {
    double v[NV][U];
#pragma unroll
    for (int k = 0; k < NV; ++k)
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const unsigned i = base + u * NT + threadIdx.x;
        v[k][u] = i < G ? partials[(size_t)k * G + i] : 0.0;
      }
#pragma unroll
    for (int k = 0; k < NV; ++k)
#pragma unroll
      for (int u = 0; u < U; ++u) tot[k] += v[k][u];
  }
  block_sum<NT, NV>(tot, sh);
}

// Scalar steps run by the finalize kernel after a reduction.
enum FinalizeOp : int {
  kFinOut = 0,        // out[k] = sum_k
  kFinPcgInit = 1,    // rho = sum0, bnorm2 = rr = sum1, reset the iteration
  kFinPcgAlpha = 2,   // pq = sum0, alpha = rho / pq, breakdown test
  kFinPcgBeta = 3,    // rho' = sum0, rr = sum1, beta, iteration count, stop test
  kFinOc = 4,         // OC bisection update from the volume sum0
  kFinReplace = 5,    // rho = sum0, rr = sum1 of the replaced (true) residual; stop test
};

// Launches the single-block finalize of a reduction over G producer blocks.
void launch_finalize(int op, int nv, const double *partials, unsigned G, PcgScalars *S,
                     OcState *O, double *out, hipStream_t s);

}  // namespace icarat_hip
