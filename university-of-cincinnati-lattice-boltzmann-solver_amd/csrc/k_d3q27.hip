// D3Q27 kernels: fused pull-stream + entropic collide on a periodic box,
// layout conversion, macroscopic fields and the seeded Taylor-Green initialiser.
//
// Layout in HBM: structure of arrays, one contiguous array per velocity,
// f[q * n + node], node = x + nx*(y + ny*z).  One thread owns one node; a
// 64-wide wavefront covers 64 consecutive x, so each of the 27 pulls
// f_q(x - c_q) is a unit-stride 256 B (FP32) wavefront load shifted by at most
// one element, and the 27 stores are aligned.  Every population of the source
// buffer is read by exactly one node, so the step moves the algorithmic
// 27 * 4 B * 2 = 216 B per node (FP32) through HBM.
//
// Reference correspondence: the pull f_q(x) = fout_q(x - c_q) is the push
// fi(x + c_q) = fout(x) of looping/timestepping/streaming.f95:365-368
// performed from the receiving side, followed by the moment, equilibrium,
// alpha and relaxation work of calculate_flow_variables.f95:134-333 and
// collision.f95:153-513 for the same node, without the intermediate fi/fout
// round trip through memory.
#include "d3q27.h"
#include "kernels.h"
#include "philox.h"

namespace uclbs {

namespace {

__device__ __forceinline__ void wave_stats(unsigned long long* __restrict__ stats, bool active,
                                           int iters, int status, bool fallback,
                                           bool logpath) {
  const unsigned long long act = __ballot(active);
  if (act == 0ull) return;
  unsigned long long it = 0;
#pragma unroll
  for (int b = 0; b < 6; ++b)
    it += (unsigned long long)__popcll(__ballot(active && ((iters >> b) & 1))) << b;
  const unsigned long long nt = __popcll(__ballot(active && status == ALPHA_TRIVIAL));
  const unsigned long long nb = __popcll(__ballot(active && status == ALPHA_BOUND));
  const unsigned long long nn = __popcll(__ballot(active && status == ALPHA_NONPOS));
  const unsigned long long nc = __popcll(__ballot(active && status == ALPHA_NOCONV));
  const unsigned long long nf = __popcll(__ballot(active && fallback));
  const unsigned long long nl = __popcll(__ballot(active && logpath));
  if (__lane_id() == 0) {
    unsigned long long* s = stats + (size_t)((blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64) &
                                             (kStatSlots - 1)) * kStatFields;
    atomicAdd(s + ST_NODES, (unsigned long long)__popcll(act));
    if (it) atomicAdd(s + ST_ITERS, it);
    if (nt) atomicAdd(s + ST_TRIVIAL, nt);
    if (nb) atomicAdd(s + ST_BOUND, nb);
    if (nn) atomicAdd(s + ST_NONPOS, nn);
    if (nc) atomicAdd(s + ST_NOCONV, nc);
    if (nf) atomicAdd(s + ST_FALLBACK, nf);
    if (nl) atomicAdd(s + ST_LOGPATH, nl);
  }
}

template <bool kShifted, bool kKeepAlpha>
__global__ __launch_bounds__(256) void k27_step_periodic(const void* __restrict__ fin_,
                                                         void* __restrict__ fout_,
                                                         double* __restrict__ alpha_out,
                                                         unsigned long long* __restrict__ stats,
                                                         Box b, StepArgs a) {
  using T = typename std::conditional<kShifted, float, double>::type;
  const T* __restrict__ fin = static_cast<const T*>(fin_);
  T* __restrict__ fout = static_cast<T*>(fout_);
  const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool active = gid < b.n;
  const int64_t node = active ? gid : 0;
  const int x = (int)(node % b.nx);
  const int64_t yz = node / b.nx;
  const int y = (int)(yz % b.ny);
  const int z = (int)(yz / b.ny);
  const int64_t sx = 1, sy = b.nx, sz = (int64_t)b.nx * b.ny;
  // source coordinate x - c for c = -1, 0, +1 (periodic wrap)
  int64_t ox[3], oy[3], oz[3];
  ox[0] = (x + 1 == b.nx ? 0 : x + 1) * sx;
  ox[1] = (int64_t)x * sx;
  ox[2] = (x == 0 ? b.nx - 1 : x - 1) * sx;
  oy[0] = (y + 1 == b.ny ? 0 : y + 1) * sy;
  oy[1] = (int64_t)y * sy;
  oy[2] = (y == 0 ? b.ny - 1 : y - 1) * sy;
  oz[0] = (z + 1 == b.nz ? 0 : z + 1) * sz;
  oz[1] = (int64_t)z * sz;
  oz[2] = (z == 0 ? b.nz - 1 : z - 1) * sz;

  T f[27];
#pragma unroll
  for (int i = 0; i < 27; ++i)
    f[i] = fin[(int64_t)i * b.n + ox[CX27[i] + 1] + oy[CY27[i] + 1] + oz[CZ27[i] + 1]];

  AlphaCfg ac{a.tol, a.maxit, a.alpha_mode, a.thr};
  double rho, u[3];
  AlphaOut ao;
  if constexpr (kShifted) ao = collide27_shifted<false>(f, a.beta, ac, 2.0, rho, u);
  else ao = collide27_fp64(f, a.beta, ac, 2.0, rho, u);

  if (active) {
    // a deferred node stores its pre-collision state; the fixup kernel collides it in place
#pragma unroll
    for (int i = 0; i < 27; ++i) fout[(int64_t)i * b.n + node] = f[i];
    if constexpr (kKeepAlpha) alpha_out[node] = ao.alpha;
  }
  if constexpr (kShifted) {
    const bool defer = active && ao.status == ALPHA_DEFER;
    const unsigned long long dm = __ballot(defer);
    if (dm) {
      unsigned base = 0;
      const int leader = __ffsll((long long)dm) - 1;
      if ((int)__lane_id() == leader) base = atomicAdd(a.defer_count + a.parity, (unsigned)__popcll(dm));
      base = __shfl(base, leader);
      if (defer) a.defer_list[base + __popcll(dm & ((1ull << __lane_id()) - 1ull))] = (unsigned)node;
    }
  }
  wave_stats(stats, active && ao.status != ALPHA_DEFER, ao.iters, ao.status, false, ao.logpath);
}

// Collides, in place, the nodes the streaming kernel queued (max|x| beyond the
// fused series order): full FP64 x_i, higher series orders or logarithms.
// Grid-stride over the device-side count; also clears the other parity's
// counter for the next step.
__global__ __launch_bounds__(256) void k27_fixup(float* __restrict__ f_, double* __restrict__ alpha_out,
                                                 unsigned long long* __restrict__ stats, Box b,
                                                 StepArgs a) {
  const unsigned count = a.defer_count[a.parity];
  if (blockIdx.x == 0 && threadIdx.x == 0) a.defer_count[1 - a.parity] = 0u;
  AlphaCfg ac{a.tol, a.maxit, a.alpha_mode, a.thr};
  const unsigned stride = gridDim.x * blockDim.x;
  const unsigned rounds = (count + stride - 1) / stride;
  for (unsigned r = 0; r < rounds; ++r) {
    const unsigned k = r * stride + blockIdx.x * blockDim.x + threadIdx.x;
    const bool active = k < count;
    const int64_t node = active ? (int64_t)a.defer_list[k] : 0;
    float f[27];
#pragma unroll
    for (int i = 0; i < 27; ++i) f[i] = f_[(int64_t)i * b.n + node];
    double rho, u[3];
    AlphaOut ao = collide27_shifted<true>(f, a.beta, ac, 2.0, rho, u);
    if (active) {
#pragma unroll
      for (int i = 0; i < 27; ++i) f_[(int64_t)i * b.n + node] = f[i];
      if (alpha_out) alpha_out[node] = ao.alpha;
    }
    wave_stats(stats, active, ao.iters, ao.status, false, ao.logpath);
  }
}

template <bool kShifted>
__global__ void k27_from_fp64(const double* __restrict__ src, void* __restrict__ dst_, Box b) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= b.n) return;
#pragma unroll
  for (int i = 0; i < 27; ++i) {
    const double v = src[(int64_t)i * b.n + k];
    if constexpr (kShifted) static_cast<float*>(dst_)[(int64_t)i * b.n + k] = (float)(v - W27[i]);
    else static_cast<double*>(dst_)[(int64_t)i * b.n + k] = v;
  }
}

template <bool kShifted>
__global__ void k27_to_fp64(const void* __restrict__ src_, double* __restrict__ dst, Box b) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= b.n) return;
#pragma unroll
  for (int i = 0; i < 27; ++i) {
    double v;
    if constexpr (kShifted) v = W27[i] + (double)static_cast<const float*>(src_)[(int64_t)i * b.n + k];
    else v = static_cast<const double*>(src_)[(int64_t)i * b.n + k];
    dst[(int64_t)i * b.n + k] = v;
  }
}

template <bool kShifted>
__global__ void k27_macro(const void* __restrict__ src_, Box b, double* __restrict__ rho,
                          double* __restrict__ ux, double* __restrict__ uy,
                          double* __restrict__ uz) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= b.n) return;
  double r = 0.0, jx = 0.0, jy = 0.0, jz = 0.0;
#pragma unroll
  for (int i = 0; i < 27; ++i) {
    double v;
    if constexpr (kShifted) v = (double)static_cast<const float*>(src_)[(int64_t)i * b.n + k];
    else v = static_cast<const double*>(src_)[(int64_t)i * b.n + k];
    r += v;
    jx += CX27[i] * v;
    jy += CY27[i] * v;
    jz += CZ27[i] * v;
  }
  if constexpr (kShifted) r += 1.0;
  rho[k] = r;
  ux[k] = jx / r;
  uy[k] = jy / r;
  uz[k] = jz / r;
}

// Taylor-Green vortex: u = U0 (sin X cos Y cos Z, -cos X sin Y cos Z, 0) plus
// uniform noise in (-amp, amp) per component from Philox4x32-10 keyed by the
// seed and counted by (node, component); f = feq(1, u).
template <bool kShifted>
__global__ void k27_init_tgv(void* __restrict__ dst_, Box b, double u0, double amp,
                             uint64_t seed) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= b.n) return;
  const int x = (int)(k % b.nx);
  const int64_t yz = k / b.nx;
  const int y = (int)(yz % b.ny), z = (int)(yz / b.ny);
  const double tp = 6.283185307179586476925286766559;
  const double X = tp * x / b.nx, Y = tp * y / b.ny, Z = tp * z / b.nz;
  double u[3];
  u[0] = u0 * sin(X) * cos(Y) * cos(Z);
  u[1] = -u0 * cos(X) * sin(Y) * cos(Z);
  u[2] = 0.0;
  for (int c = 0; c < 3; ++c) u[c] += amp * (2.0 * philox_uniform((uint64_t)k, (uint32_t)c, seed) - 1.0);
  double px[3], py[3], pz[3];
  auto psi = [](double uu, double* p) {
    const double s = sqrt(1.0 + 3.0 * uu * uu);
    const double bb = 2.0 - s;
    p[1] = bb;
    p[2] = bb * (2.0 * uu + s) / (1.0 - uu);
    p[0] = bb * (1.0 - uu) / (2.0 * uu + s);
  };
  psi(u[0], px); psi(u[1], py); psi(u[2], pz);
#pragma unroll
  for (int i = 0; i < 27; ++i) {
    const double feq = W27[i] * px[CX27[i] + 1] * py[CY27[i] + 1] * pz[CZ27[i] + 1];
    if constexpr (kShifted) static_cast<float*>(dst_)[(int64_t)i * b.n + k] = (float)(feq - W27[i]);
    else static_cast<double*>(dst_)[(int64_t)i * b.n + k] = feq;
  }
}

inline unsigned blocks_for(int64_t n, int t) { return (unsigned)((n + t - 1) / t); }

}  // namespace

hipError_t launch_d3q27_step_periodic(const Box& b, const void* fin, void* fout, double* alpha,
                                      unsigned long long* stats, const StepArgs& a,
                                      hipStream_t s) {
  const unsigned g = blocks_for(b.n, 256);
  if (a.shifted) {
    if (alpha) k27_step_periodic<true, true><<<g, 256, 0, s>>>(fin, fout, alpha, stats, b, a);
    else k27_step_periodic<true, false><<<g, 256, 0, s>>>(fin, fout, alpha, stats, b, a);
  } else {
    if (alpha) k27_step_periodic<false, true><<<g, 256, 0, s>>>(fin, fout, alpha, stats, b, a);
    else k27_step_periodic<false, false><<<g, 256, 0, s>>>(fin, fout, alpha, stats, b, a);
  }
  return hipGetLastError();
}

hipError_t launch_d3q27_fixup(const Box& b, void* fout, double* alpha,
                              unsigned long long* stats, const StepArgs& a, hipStream_t s) {
  k27_fixup<<<512, 256, 0, s>>>(static_cast<float*>(fout), alpha, stats, b, a);
  return hipGetLastError();
}

hipError_t launch_d3q27_from_fp64(const Box& b, const double* src, void* dst, int shifted,
                                  hipStream_t s) {
  if (shifted) k27_from_fp64<true><<<blocks_for(b.n, 256), 256, 0, s>>>(src, dst, b);
  else k27_from_fp64<false><<<blocks_for(b.n, 256), 256, 0, s>>>(src, dst, b);
  return hipGetLastError();
}

hipError_t launch_d3q27_to_fp64(const Box& b, const void* src, double* dst, int shifted,
                                hipStream_t s) {
  if (shifted) k27_to_fp64<true><<<blocks_for(b.n, 256), 256, 0, s>>>(src, dst, b);
  else k27_to_fp64<false><<<blocks_for(b.n, 256), 256, 0, s>>>(src, dst, b);
  return hipGetLastError();
}

hipError_t launch_d3q27_macro(const Box& b, const void* f, int shifted, double* rho,
                              double* ux, double* uy, double* uz, hipStream_t s) {
  if (shifted) k27_macro<true><<<blocks_for(b.n, 256), 256, 0, s>>>(f, b, rho, ux, uy, uz);
  else k27_macro<false><<<blocks_for(b.n, 256), 256, 0, s>>>(f, b, rho, ux, uy, uz);
  return hipGetLastError();
}

hipError_t launch_d3q27_init_tgv(const Box& b, void* f, int shifted, double u0,
                                 double noise_amp, uint64_t seed, hipStream_t s) {
  if (shifted) k27_init_tgv<true><<<blocks_for(b.n, 256), 256, 0, s>>>(f, b, u0, noise_amp, seed);
  else k27_init_tgv<false><<<blocks_for(b.n, 256), 256, 0, s>>>(f, b, u0, noise_amp, seed);
  return hipGetLastError();
}

}  // namespace uclbs
