// Plain device-to-device copy used as the bandwidth yardstick next to the
// vendor HBM peak, and as the known-size calibration for the HBM byte counters.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "../../include/uclbs_hip.h"

namespace {

// one 16-byte element per thread over a grid that covers the array: every
// wave issues its load at once (a grid-stride loop keeps fewer loads in
// flight and measured 4.5-5.5 TB/s against 6.2 TB/s for this form)
__global__ __launch_bounds__(256) void k_copy4(const float4* __restrict__ a, float4* __restrict__ b,
                                               int64_t n4) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n4) b[k] = a[k];
}

}  // namespace

extern "C" int uclbs_copy_bandwidth(int64_t nbytes, int reps, double* gbps) {
  if (nbytes < 4096 || nbytes > ((int64_t)1 << 36) || reps <= 0 || reps > 1000 || !gbps)
    return UCLBS_E_ARG;
  const int64_t n4 = nbytes / 16;
  float4 *a = nullptr, *b = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipError_t e = hipMalloc(&a, n4 * 16);
  if (e == hipSuccess) e = hipMalloc(&b, n4 * 16);
  if (e == hipSuccess) e = hipMemset(a, 0, n4 * 16);
  if (e == hipSuccess) e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  const unsigned grid = (unsigned)((n4 + 255) / 256);
  if (e == hipSuccess) {
    k_copy4<<<grid, 256>>>(a, b, n4);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipEventRecord(e0, nullptr);
  for (int r = 0; r < reps && e == hipSuccess; ++r) {
    k_copy4<<<grid, 256>>>(a, b, n4);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipEventRecord(e1, nullptr);
  if (e == hipSuccess) e = hipEventSynchronize(e1);
  float ms = 0.f;
  if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
  if (e == hipSuccess) *gbps = 2.0 * (double)(n4 * 16) * reps / (ms * 1e-3) / 1e9;
  if (a) (void)hipFree(a);
  if (b) (void)hipFree(b);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  return e == hipSuccess ? UCLBS_OK : UCLBS_E_HIP;
}
