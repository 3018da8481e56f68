// noalloc.hip -- the library's one entry point that allocates nothing, for
// callers that cannot call hipMalloc: code under HIP graph capture, tight
// serving loops, harnesses that forbid allocation.
//
// acu_prep_mirror_error_noalloc is acu_prep_mirror_error_dev (the mirror
// error of the centre-of-rotation search, tomography/
// projections_to_sinogram_hpc.F90:260-283) with its FP64 block partials in
// `scratch`, a file-scope device array of NOALLOC_SCRATCH_DOUBLES doubles,
// instead of a buffer allocated per call, and its result written to device
// memory on the context's stream without a wait.  The kernels are the same
// (prep.hip: mirror_kernel writes every partial it owns, sum_kernel reads
// exactly those), so nothing in the scratch has to start at zero and the
// result is bit-identical to the allocating entry point's.
//
// Each GPU holds its own copy of `scratch` at its own address.  In host code
// the name `scratch` is only a host-side stand-in, so the device address is
// looked up once per device with hipGetSymbolAddress (noalloc_setup, called
// by acu_ctx_create with the device current) and only that address is ever
// passed to kernels.  Every call on one GPU shares that GPU's array: calls
// that overlap on one GPU (other streams, other threads) race.
#include "acu_prep.hpp"
#include "acumen.h"

#include <mutex>
#include <string>

namespace {

constexpr int NOALLOC_SCRATCH_DOUBLES = ACU_NOALLOC_SCRATCH_DOUBLES;  // 512 KiB per GPU

__device__ __align__(256) double scratch[NOALLOC_SCRATCH_DOUBLES];

constexpr int MAX_DEVICES = 64;
std::mutex g_mu;
double* g_scratch[MAX_DEVICES] = {};

}  // namespace

namespace acu {

// Device address of this GPU's scratch; `device` must be current.
void noalloc_setup(int device) {
  if (device < 0 || device >= MAX_DEVICES) throw Error("noalloc_setup: device index out of range");
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_scratch[device]) return;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p)
    throw Error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
  g_scratch[device] = static_cast<double*>(p);
}

}  // namespace acu

extern "C" int acu_prep_mirror_error_noalloc(acu_ctx* ctx, int n, int add_nx, float theta, float shift,
                                             const float* p0_dev, const float* p180_dev, double* err_dev) {
  if (!ctx || n < 1 || (add_nx != 0 && add_nx != 1) || !p0_dev || !p180_dev || !err_dev)
    return acu::record_error(ACU_ERR_ARG, "acu_prep_mirror_error_noalloc: bad argument");
  const acu::PrepGeom g = acu::prep_geom(n, add_nx, theta, shift);
  const long long nb = acu::mirror_blocks(g);
  if (nb > NOALLOC_SCRATCH_DOUBLES)
    return acu::record_error(ACU_ERR_ARG, "acu_prep_mirror_error_noalloc: " + std::to_string(nb) +
                                              " block partials exceed the scratch of " +
                                              std::to_string(NOALLOC_SCRATCH_DOUBLES) +
                                              " (n at most 4095)");
  const int dev = acu::ctx_device(ctx);
  double* part = (dev >= 0 && dev < MAX_DEVICES) ? g_scratch[dev] : nullptr;
  if (!part) return acu::record_error(ACU_ERR_STATE, "acu_prep_mirror_error_noalloc: scratch not set up");
  hipError_t e = hipSetDevice(dev);
  if (e == hipSuccess) {
    acu::launch_mirror(g, p0_dev, p180_dev, part, err_dev, acu::ctx_stream(ctx));
    e = hipGetLastError();
  }
  if (e != hipSuccess)
    return acu::record_error(ACU_ERR_HIP, std::string("acu_prep_mirror_error_noalloc: ") + hipGetErrorString(e));
  return ACU_OK;
}
