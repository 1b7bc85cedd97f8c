// noalloc.hip -- element stresses without a device allocation.
//
// icarat_hip_mises stages the nelx x 6 element stresses in a buffer it
// allocates for the call.  The variant here stages them in `scratch`, a
// file-scope device array of fixed size, so the call allocates nothing: for
// callers that cannot allocate device memory (code under graph capture, tight
// serving loops).  Every call on one GPU shares that GPU's array, which keeps
// the last call's stresses; calls that overlap on one GPU race.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "context.hpp"
#include "kernels.hpp"

namespace icarat_hip {

namespace {

// 6 doubles per element for up to 2^20 elements: 48 MiB per GPU
constexpr long long kNoallocMaxElements = 1LL << 20;
constexpr size_t kScratchDoubles = 6 * (size_t)kNoallocMaxElements;

static __device__ __align__(256) double scratch[kScratchDoubles];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
double *g_scratch[kMaxDevices] = {};

}  // namespace

long long noalloc_max_elements() { return kNoallocMaxElements; }

// Device address of this GPU's copy of `scratch`, looked up once per device
// with that device current (called from the context's device set-up).  The
// host-side name of the array is not a device address: kernels, copies and
// fills take this pointer.
void noalloc_init_device(int device) {
  if (device < 0 || device >= kMaxDevices) throw Error(kInternal, "device index out of range");
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_scratch[device]) return;
  void *p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p)
    throw Error(kHipError, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
  g_scratch[device] = static_cast<double *>(p);
}

static double *scratch_of(int device) {
  std::lock_guard<std::mutex> lk(g_mu);
  double *p = device >= 0 && device < kMaxDevices ? g_scratch[device] : nullptr;
  if (!p) throw Error(kInternal, "no-allocation scratch of this device is not set up");
  return p;
}

void Context::mises_noalloc(double *mises, double *stress) {
  if (stress && L_.nelx > kNoallocMaxElements)
    throw Error(kInvalidArgument, "icarat_hip_mises_noalloc: " + std::to_string(L_.nelx) +
                                      " elements exceed the maximum of " +
                                      std::to_string(kNoallocMaxElements) +
                                      " for the element stresses");
  // a solution of the current density implies the device set-up, numbering,
  // loads and densities are in place (no set-up step, which would allocate)
  if (!dev_ready_ || !density_ready_ || !solved_ || fext_dirty_)
    throw Error(kNotReady, "no solution for the current density");
  const auto &g = prob_.spec();
  const double h[3] = {g.lx / g.nx, g.ly / g.ny, g.lz / g.nz};
  double *st = stress ? scratch_of(device_) : nullptr;
  launch_flush_small(n3_, d_u, d_nv, stream_);
  launch_mises64(L_, d_dens, d_nv, g.young1, g.young2, g.penalty, g.poisson, h, d_e[6], st,
                 stream_);
  HIPCHK(hipMemcpyAsync(mises, d_e[6].p, sizeof(double) * L_.nelx, hipMemcpyDeviceToHost, stream_));
  if (stress)
    HIPCHK(hipMemcpyAsync(stress, st, sizeof(double) * 6 * L_.nelx, hipMemcpyDeviceToHost,
                          stream_));
  sync();
}

}  // namespace icarat_hip
