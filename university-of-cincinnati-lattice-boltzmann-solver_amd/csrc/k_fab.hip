// AMReX patch data (FArrayBox) on the device <-> a level's dense node order.
//
// The reference keeps a level's populations, second populations and
// multipliers in MultiFabs (amr_info_holder.f95:50-76): per box an FP64 array
// of ncomp components, component-major, each component the nodal box grown by
// the ghost width, x fastest.  These kernels gather the valid region of such
// an array into [ncomp][n] (node = a + nx (b + ny c)) and scatter it back; the
// ghost cells are neither read nor written.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "kernels.h"

namespace uclbs {
namespace {

__global__ __launch_bounds__(256) void k_fab_gather(const double* __restrict__ fab, int ghost,
                                                    int64_t cstride, Box b, int ncomp,
                                                    double* __restrict__ dense) {
  const int64_t n = b.n, tot = n * ncomp;
  const int64_t gx = b.nx + 2 * ghost, gy = b.ny + 2 * ghost;
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < tot;
       k += (int64_t)gridDim.x * blockDim.x) {
    const int64_t comp = k / n, node = k - comp * n;
    const int64_t a = node % b.nx, yz = node / b.nx, bb = yz % b.ny, c = yz / b.ny;
    dense[k] = fab[comp * cstride + (a + ghost) + gx * ((bb + ghost) + gy * (c + ghost))];
  }
}

__global__ __launch_bounds__(256) void k_fab_scatter(const double* __restrict__ dense, Box b,
                                                     int ncomp, int ghost, int64_t cstride,
                                                     double* __restrict__ fab) {
  const int64_t n = b.n, tot = n * ncomp;
  const int64_t gx = b.nx + 2 * ghost, gy = b.ny + 2 * ghost;
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < tot;
       k += (int64_t)gridDim.x * blockDim.x) {
    const int64_t comp = k / n, node = k - comp * n;
    const int64_t a = node % b.nx, yz = node / b.nx, bb = yz % b.ny, c = yz / b.ny;
    fab[comp * cstride + (a + ghost) + gx * ((bb + ghost) + gy * (c + ghost))] = dense[k];
  }
}

unsigned fab_grid(const Box& b, int ncomp) {
  const int64_t blocks = (b.n * ncomp + 255) / 256;
  return (unsigned)(blocks < 65536 ? blocks : 65536);
}

}  // namespace

hipError_t launch_fab_gather(const double* fab, int ghost, int64_t cstride, const Box& b,
                             int ncomp, double* dense, hipStream_t s) {
  hipLaunchKernelGGL(k_fab_gather, dim3(fab_grid(b, ncomp)), dim3(256), 0, s, fab, ghost, cstride,
                     b, ncomp, dense);
  return hipGetLastError();
}

hipError_t launch_fab_scatter(const double* dense, const Box& b, int ncomp, int ghost,
                              int64_t cstride, double* fab, hipStream_t s) {
  hipLaunchKernelGGL(k_fab_scatter, dim3(fab_grid(b, ncomp)), dim3(256), 0, s, dense, b, ncomp,
                     ghost, cstride, fab);
  return hipGetLastError();
}

}  // namespace uclbs
