// Coarse-fine exchange between nested D3Q27 levels (refinement ratio 2 in
// space and time, nodal grids).
//
// Reference: looping/timestepping/interlevel/interlevel_interpolation.f95:
// 114-493 fills the fine ghost nodes with amrex_fillcoarsepatch, i.e. the
// coarse populations interpolated bilinearly per axis on the nodal grid
// (amrex_interp_node_bilinear) and linearly in time between the coarse
// level's previous and current states; a negative or NaN interpolated
// population is replaced by the coarse equilibrium (:417-431).
// fine_to_coarse.f95:25-65 averages the fine level down onto the covered
// coarse nodes with amrex_average_down_nodes, which for nodal data at ratio 2
// is injection from the coincident fine node.
//
// A level l >= 1 is stored with a one-node ghost shell around its valid box;
// node (a, b, c) of the stored box has the level's global index
// origin + (a, b, c).  Fine global index G sits at coarse coordinate G / 2:
// on a coincident node (G even) the coarse value is taken, between two coarse
// nodes (G odd) their mean, per axis (tensor product: 1, 2, 4 or 8 coarse
// nodes).  Populations are interpolated in their storage form (the FP32
// shift by the lattice weight is the same on every level).
#include "kernels.h"
#include "lattice.h"

#include <type_traits>

namespace uclbs {

namespace {

template <bool kShifted>
__device__ __forceinline__ void repair_store27(double (&f)[27],
                                               typename std::conditional<kShifted, float, double>::type* __restrict__ ff,
                                               unsigned fnn, unsigned node,
                                               unsigned long long* __restrict__ stats,
                                               unsigned mask = kAllPops27);

template <bool kShifted>
__global__ __launch_bounds__(256) void k27_coarse_to_fine(
    const void* __restrict__ cprev_, const void* __restrict__ cnew_, LevelMap cm,
    void* __restrict__ ff_, LevelMap fm, const unsigned* __restrict__ ghosts, unsigned nghost,
    double theta, unsigned long long* __restrict__ stats) {
  using T = typename std::conditional<kShifted, float, double>::type;
  const T* __restrict__ cprev = static_cast<const T*>(cprev_);
  const T* __restrict__ cnew = static_cast<const T*>(cnew_);
  T* __restrict__ ff = static_cast<T*>(ff_);
  const unsigned j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nghost) return;
  const unsigned node = ghosts ? ghosts[j] : j;  // no list: every stored node
  const unsigned fnx = (unsigned)fm.box.nx, fny = (unsigned)fm.box.ny;
  const int loc[3] = {(int)(node % fnx), (int)((node / fnx) % fny), (int)(node / (fnx * fny))};
  const int cn[3] = {cm.box.nx, cm.box.ny, cm.box.nz};
  int ci[3][2];
  double wt[3][2];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const int G = fm.origin[d] + loc[d];
    const int odd = G & 1;
    int c0 = (G >> 1) - cm.origin[d];
    int c1 = c0 + odd;
    if (cm.per[d]) {
      c0 = (c0 % cn[d] + cn[d]) % cn[d];
      c1 = (c1 % cn[d] + cn[d]) % cn[d];
    }
    ci[d][0] = c0;
    ci[d][1] = c1;
    wt[d][0] = odd ? 0.5 : 1.0;
    wt[d][1] = odd ? 0.5 : 0.0;
  }
  const unsigned cnn = (unsigned)cm.box.n, fnn = (unsigned)fm.box.n;
  double f[27];
#pragma unroll
  for (int i = 0; i < 27; ++i) f[i] = 0.0;
  for (int cz = 0; cz < 2; ++cz)
    for (int cy = 0; cy < 2; ++cy)
      for (int cx = 0; cx < 2; ++cx) {
        const double w = wt[0][cx] * wt[1][cy] * wt[2][cz];
        if (w == 0.0) continue;
        const unsigned k = (unsigned)ci[0][cx] +
                           (unsigned)cm.box.nx * ((unsigned)ci[1][cy] + (unsigned)cm.box.ny * (unsigned)ci[2][cz]);
#pragma unroll
        for (int i = 0; i < 27; ++i) {
          const double a = (double)cprev[(size_t)i * cnn + k], b = (double)cnew[(size_t)i * cnn + k];
          f[i] += w * ((1.0 - theta) * a + theta * b);
        }
      }
  repair_store27<kShifted>(f, ff, fnn, node, stats);
}

// negative (or NaN) interpolated populations take the equilibrium of the
// interpolated moments; then the node is stored
template <bool kShifted>
__device__ __forceinline__ void repair_store27(double (&f)[27],
                                               typename std::conditional<kShifted, float, double>::type* __restrict__ ff,
                                               unsigned fnn, unsigned node,
                                               unsigned long long* __restrict__ stats,
                                               unsigned mask) {
  using T = typename std::conditional<kShifted, float, double>::type;
  bool neg = false;
#pragma unroll
  for (int i = 0; i < 27; ++i) neg |= !((kShifted ? W27[i] + f[i] : f[i]) > 0.0);
  if (neg) {
    double rho = 0.0, jx = 0.0, jy = 0.0, jz = 0.0;
#pragma unroll
    for (int i = 0; i < 27; ++i) {
      const double fi = kShifted ? W27[i] + f[i] : f[i];
      rho += fi;
      jx += CX27[i] * fi;
      jy += CY27[i] * fi;
      jz += CZ27[i] * fi;
    }
    const double u[3] = {jx / rho, jy / rho, jz / rho};
    double p[3][3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const double s = sqrt(1.0 + 3.0 * u[d] * u[d]), bb = 2.0 - s;
      p[d][1] = bb;
      p[d][2] = bb * (2.0 * u[d] + s) / (1.0 - u[d]);
      p[d][0] = bb * (1.0 - u[d]) / (2.0 * u[d] + s);
    }
    unsigned long long nrep = 0;
#pragma unroll
    for (int i = 0; i < 27; ++i) {
      const double fi = kShifted ? W27[i] + f[i] : f[i];
      if (!(fi > 0.0)) {
        const double e = rho * W27[i] * p[0][CX27[i] + 1] * p[1][CY27[i] + 1] * p[2][CZ27[i] + 1];
        f[i] = kShifted ? e - W27[i] : e;
        ++nrep;
      }
    }
    atomicAdd(stats + (size_t)(blockIdx.x & (kStatSlots - 1)) * kStatFields + ST_REPAIR, nrep);
  }
#pragma unroll
  for (int i = 0; i < 27; ++i)
    if ((mask >> i) & 1u) ff[(size_t)i * fnn + node] = (T)f[i];
}

// ghost nodes of a level of many boxes from precomputed coarse stencils; the
// same arithmetic and order as k27_coarse_to_fine
template <bool kShifted>
__global__ __launch_bounds__(256) void k27_c2f_stencils(const void* __restrict__ cprev_,
                                                        const void* __restrict__ cnew_, unsigned cnn,
                                                        void* __restrict__ ff_, unsigned fnn,
                                                        const C2FStencil* __restrict__ st,
                                                        unsigned nst, double theta,
                                                        unsigned long long* __restrict__ stats,
                                                        unsigned extra_mask) {
  using T = typename std::conditional<kShifted, float, double>::type;
  const T* __restrict__ cprev = static_cast<const T*>(cprev_);
  const T* __restrict__ cnew = static_cast<const T*>(cnew_);
  const unsigned j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nst) return;
  const C2FStencil s = st[j];
  double f[27];
#pragma unroll
  for (int i = 0; i < 27; ++i) f[i] = 0.0;
  for (int q = 0; q < 8; ++q) {
    const double w = (double)s.w[q];
    if (w == 0.0) continue;
    const unsigned k = s.src[q];
    if (theta == 0.0) {
      // (1 - 0) a + 0 b is a: the current state is not read
#pragma unroll
      for (int i = 0; i < 27; ++i) f[i] += w * (double)cprev[(size_t)i * cnn + k];
    } else {
#pragma unroll
      for (int i = 0; i < 27; ++i) {
        const double a = (double)cprev[(size_t)i * cnn + k], b = (double)cnew[(size_t)i * cnn + k];
        f[i] += w * ((1.0 - theta) * a + theta * b);
      }
    }
  }
  repair_store27<kShifted>(f, static_cast<T*>(ff_), fnn, s.dst, stats, s.mask | extra_mask);
}

// src and dst may be the same level (ghost copies): a pair's source is a valid
// node, never a ghost, so no pair reads what another writes
template <typename T>
__global__ __launch_bounds__(256) void k27_copy_pairs(const T* src, unsigned snn, T* dst, unsigned dnn,
                                                      const uint3* __restrict__ pairs,
                                                      unsigned npairs) {
  const unsigned j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= npairs) return;
  const uint3 p = pairs[j];
#pragma unroll
  for (int i = 0; i < 27; ++i)
    if ((p.z >> i) & 1u) dst[(size_t)i * dnn + p.x] = src[(size_t)i * snn + p.y];
}

// injection: covered coarse node C <- fine node 2C
template <typename T>
__global__ __launch_bounds__(256) void k27_fine_to_coarse(const T* __restrict__ ff, LevelMap fm,
                                                          T* __restrict__ fc, LevelMap cm,
                                                          Range3 cover) {
  const unsigned ex = (unsigned)(cover.hi[0] - cover.lo[0] + 1),
                 ey = (unsigned)(cover.hi[1] - cover.lo[1] + 1),
                 ez = (unsigned)(cover.hi[2] - cover.lo[2] + 1);
  const unsigned t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ex * ey * ez) return;
  const int C[3] = {cover.lo[0] + (int)(t % ex), cover.lo[1] + (int)((t / ex) % ey),
                    cover.lo[2] + (int)(t / (ex * ey))};
  const unsigned kc = (unsigned)(C[0] - cm.origin[0]) +
                      (unsigned)cm.box.nx * ((unsigned)(C[1] - cm.origin[1]) +
                                             (unsigned)cm.box.ny * (unsigned)(C[2] - cm.origin[2]));
  const unsigned kf = (unsigned)(2 * C[0] - fm.origin[0]) +
                      (unsigned)fm.box.nx * ((unsigned)(2 * C[1] - fm.origin[1]) +
                                             (unsigned)fm.box.ny * (unsigned)(2 * C[2] - fm.origin[2]));
  const unsigned cnn = (unsigned)cm.box.n, fnn = (unsigned)fm.box.n;
#pragma unroll
  for (int i = 0; i < 27; ++i) fc[(size_t)i * cnn + kc] = ff[(size_t)i * fnn + kf];
}

}  // namespace

hipError_t launch_coarse_to_fine27(const void* cprev, const void* cnew, const LevelMap& cm, void* ff,
                                   const LevelMap& fm, const unsigned* ghosts, unsigned nghost,
                                   double theta, int shifted, unsigned long long* stats,
                                   hipStream_t s) {
  if (!nghost) return hipSuccess;
  const unsigned g = (nghost + 255) / 256;
  if (shifted)
    k27_coarse_to_fine<true><<<g, 256, 0, s>>>(cprev, cnew, cm, ff, fm, ghosts, nghost, theta, stats);
  else
    k27_coarse_to_fine<false><<<g, 256, 0, s>>>(cprev, cnew, cm, ff, fm, ghosts, nghost, theta, stats);
  return hipGetLastError();
}

hipError_t launch_copy_pairs27(const void* src, unsigned nn_src, void* dst, unsigned nn_dst,
                               const uint3* pairs, unsigned npairs, int shifted, hipStream_t s) {
  if (!npairs) return hipSuccess;
  const unsigned g = (npairs + 255) / 256;
  if (shifted)
    k27_copy_pairs<float><<<g, 256, 0, s>>>(static_cast<const float*>(src), nn_src,
                                            static_cast<float*>(dst), nn_dst, pairs, npairs);
  else
    k27_copy_pairs<double><<<g, 256, 0, s>>>(static_cast<const double*>(src), nn_src,
                                             static_cast<double*>(dst), nn_dst, pairs, npairs);
  return hipGetLastError();
}

hipError_t launch_c2f_stencils27(const void* cprev, const void* cnew, unsigned nn_c, void* ff,
                                 unsigned nn_f, const C2FStencil* st, unsigned nst, double theta,
                                 int shifted, unsigned long long* stats, hipStream_t s,
                                 int all_pops) {
  if (!nst) return hipSuccess;
  const unsigned g = (nst + 255) / 256;
  if (shifted)
    k27_c2f_stencils<true><<<g, 256, 0, s>>>(cprev, cnew, nn_c, ff, nn_f, st, nst, theta, stats,
                                                             all_pops ? kAllPops27 : 0u);
  else
    k27_c2f_stencils<false><<<g, 256, 0, s>>>(cprev, cnew, nn_c, ff, nn_f, st, nst, theta, stats,
                                                             all_pops ? kAllPops27 : 0u);
  return hipGetLastError();
}

hipError_t launch_fine_to_coarse27(const void* ff, const LevelMap& fm, void* fc, const LevelMap& cm,
                                   const Range3& cover, int shifted, hipStream_t s) {
  const long long n = (long long)(cover.hi[0] - cover.lo[0] + 1) * (cover.hi[1] - cover.lo[1] + 1) *
                      (cover.hi[2] - cover.lo[2] + 1);
  if (n <= 0) return hipSuccess;
  const unsigned g = (unsigned)((n + 255) / 256);
  if (shifted)
    k27_fine_to_coarse<float><<<g, 256, 0, s>>>(static_cast<const float*>(ff), fm,
                                               static_cast<float*>(fc), cm, cover);
  else
    k27_fine_to_coarse<double><<<g, 256, 0, s>>>(static_cast<const double*>(ff), fm,
                                                static_cast<double*>(fc), cm, cover);
  return hipGetLastError();
}

}  // namespace uclbs
