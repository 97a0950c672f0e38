// Calibration kernel for the VALU side of the roofline (no counterpart in the reference).
//
// The counting kernels are integer VALU-issue bound. The issue peak of gfx950 for 32-bit integer
// VALU instructions is not a published figure, so the bench measures it: this kernel keeps every
// SIMD busy with independent integer chains (xor, shift, add: no loads, no LDS, no multiplies), and
// its SQ_INSTS_VALU reading per cycle under rocprofv3 is the rate the counting kernels' own reading
// is divided by.
#include <hip/hip_runtime.h>

#include "hm_internal.hpp"

namespace hm {
namespace {

constexpr int kChains = 8;

// Lanes at or above `lanes` (of each wave's 64) skip the chains: with lanes < 64 every VALU
// instruction of the loop issues with part of EXEC cleared.
__global__ __launch_bounds__(256) void k_probe_valu(uint32_t* out, int32_t iters, uint32_t seed, int32_t lanes) {
  uint32_t x[kChains];
#pragma unroll
  for (int c = 0; c < kChains; ++c) x[c] = seed + threadIdx.x * 131u + (uint32_t)c * 977u + blockIdx.x;
  if ((int32_t)(threadIdx.x & 63) < lanes) {
    for (int32_t it = 0; it < iters; ++it) {
#pragma unroll
      for (int u = 0; u < 4; ++u) {
#pragma unroll
        for (int c = 0; c < kChains; ++c) x[c] = (x[c] ^ (0x9E3779B9u + (uint32_t)c)) + (x[c] >> 5);
      }
    }
  }
  uint32_t s = 0;
#pragma unroll
  for (int c = 0; c < kChains; ++c) s ^= x[c];
  out[(size_t)blockIdx.x * blockDim.x + threadIdx.x] = s;
}

}  // namespace
}  // namespace hm

extern "C" hm_status hm_probe_valu(int32_t iters, int32_t lanes, double* ms) {
  hm::clear_error();
  if (iters <= 0 || lanes < 1 || lanes > 64 || !ms)
    return hm::fail(HM_ERR_BAD_INPUT, "hm_probe_valu: need iters > 0, 1 <= lanes <= 64 and an output");
  int dev = hm::current_device();
  HM_HIP(hipSetDevice(dev));
  int cus = 0;
  HM_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  // 8 workgroups of 4 waves per CU: 8 waves per SIMD, enough to hide the chains' latency
  const unsigned grid = (unsigned)cus * 8u, block = 256u;
  uint32_t* out = nullptr;
  {
    hm::AllocGuard guard;
    HM_HIP(hipMalloc(&out, sizeof(uint32_t) * (size_t)grid * block));
  }
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hm_status st = HM_OK;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) st = HM_ERR_DEVICE;
  if (st == HM_OK) {
    hm::k_probe_valu<<<grid, block>>>(out, iters, 1u, lanes);  // warm-up (code object load)
    (void)hipEventRecord(e0, nullptr);
    hm::k_probe_valu<<<grid, block>>>(out, iters, 2u, lanes);
    (void)hipEventRecord(e1, nullptr);
    if (hipGetLastError() != hipSuccess || hipEventSynchronize(e1) != hipSuccess) {
      st = hm::fail(HM_ERR_DEVICE, "hm_probe_valu: launch failed");
    } else {
      float f = 0.f;
      (void)hipEventElapsedTime(&f, e0, e1);
      *ms = f;
    }
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  {
    hm::AllocGuard guard;
    (void)hipFree(out);
  }
  return st;
}
