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

// Plot3D IBLANK of the valid region from the node codes (plot3d_output.f95:
// 275-281): 1 for a fluid node (code >= 0), -1 for code -666, 0 otherwise; 1
// everywhere when the level has no link table (code null)
__global__ __launch_bounds__(256) void k_blank_gather(const int32_t* __restrict__ code, int ghost,
                                                      Box b, int32_t* __restrict__ blank) {
  const int64_t gx = b.nx + 2 * ghost, gy = b.ny + 2 * ghost;
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < b.n;
       k += (int64_t)gridDim.x * blockDim.x) {
    int32_t v = 1;
    if (code) {
      const int64_t a = k % b.nx, yz = k / b.nx, bb = yz % b.ny, c = yz / b.ny;
      const int32_t s = code[(a + ghost) + gx * ((bb + ghost) + gy * (c + ghost))];
      v = s >= 0 ? 1 : (s == -666 ? -1 : 0);
    }
    blank[k] = v;
  }
}

__global__ __launch_bounds__(256) void k_fill(double* __restrict__ p, int64_t n, double v) {
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n;
       k += (int64_t)gridDim.x * blockDim.x)
    p[k] = v;
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

hipError_t launch_blank_gather(const int32_t* code, int ghost, const Box& b, int32_t* blank,
                               hipStream_t s) {
  hipLaunchKernelGGL(k_blank_gather, dim3(fab_grid(b, 1)), dim3(256), 0, s, code, ghost, b, blank);
  return hipGetLastError();
}

hipError_t launch_fill(double* p, int64_t n, double v, hipStream_t s) {
  Box b{};
  b.n = n;
  hipLaunchKernelGGL(k_fill, dim3(fab_grid(b, 1)), dim3(256), 0, s, p, n, v);
  return hipGetLastError();
}

}  // namespace uclbs
