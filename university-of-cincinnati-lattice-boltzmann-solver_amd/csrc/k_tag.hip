// Refinement criterion on the device: the gradient of the gradient magnitude
// (GGM) of a flow quantity, as the reference tags nodes for regridding
// (GGM/ggm_calculator.f95:47-68 -> populate_q.f95:23-56,
// calculate_grad_mag.f95:33-61, calculate_gradgrad_mag.f95:37-75), on a
// one-box D3Q27 level:
//   q    the quantity per node (1 density, 2 u, 3 v, 4 w, 7 velocity
//        magnitude) from the current populations,
//   gm   = |(dq/dx, dq/dy, dq/dz)|, 0 on nodes whose own code is < 0 other
//        than -666,
//   ggm  = |(dgm/dx, dgm/dy, dgm/dz)|, 0 on nodes whose own code is < 0,
//   tag  where ggm >= limit, on the level's valid nodes.
// Each derivative is the reference's five-point "windage" stencil
// (derivatives.f95:644-700) over (x, x+1, x+2, x-1, x-2) along the axis,
// central, or one-sided by the sign of the node's velocity component.  A
// stencil node is usable when the node's own link code for it is >= 0: the
// unit links 1+a / 4+a and, for x +- 2, the codes of directions 15+a / 18+a
// (calculate_grad_mag.f95:42-55).  Those indices are the +-2 links of the
// D3Q39 layout; on a D3Q27 table they name diagonal links, and the reference
// reads them all the same, so this kernel does too.  A node outside the box
// (beyond a non-periodic face) is never usable and reads as 0; values of
// stencil nodes that are not usable are still read where the reference's
// formulas read them (the backward formula under the forward test).
// Levels without a link table: every in-box node is usable.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "kernels.h"
#include "lattice.h"

namespace uclbs {
namespace {

constexpr int kSchemeCentral = 0, kSchemeUpwind = 1, kSchemeDownwind = 2;

// derivative_windage_firstord (derivatives.f95:589-642) on (q(x), q(x+1),
// q(x-1)) with the usability of the x + 1 and x - 1 nodes
__device__ double windage1(double q0, double qp, double qm, bool okp, bool okm, double wind,
                           int scheme) {
  if (scheme == kSchemeUpwind) {
    if (wind > 0.0) return okm ? qm - q0 : 0.0;
    return okp ? q0 - qp : 0.0;
  }
  if (scheme == kSchemeDownwind) {
    if (wind < 0.0) return okm ? qm - q0 : 0.0;
    return okp ? q0 - qp : 0.0;
  }
  return 0.0;
}

// derivative_windage_secondord (derivatives.f95:644-700): q = (x, x+1,
// x+2, x-1, x-2), ok the usability of the last four
__device__ double windage2(const double q[5], const bool ok[5], double wind, int scheme) {
  if (scheme == kSchemeCentral) {
    if (ok[1] && ok[3]) return (q[3] - q[1]) / 2.0;
    if (ok[1]) return q[1] - q[0];
    if (ok[3]) return q[0] - q[3];
    return 0.0;
  }
  const bool back = scheme == kSchemeUpwind ? wind > 0.0 : wind < 0.0;
  if (back) {
    if (ok[3] && ok[4]) return (3 * q[0] - 4 * q[3] + q[4]) / 2;
  } else {
    // the forward nodes' usability gates the backward formula, as the
    // reference writes it
    if (ok[1] && ok[2]) return (-3 * q[0] + 4 * q[3] - q[4]) / 2;
  }
  return windage1(q[0], q[1], q[3], ok[1], ok[3], wind, scheme);
}

// neighbour of node (x, y, z) at offset o along axis ax (periodic axes wrap;
// -1 beyond another face)
__device__ long long nb_node(const Box& b, int x, int y, int z, int ax, int o) {
  int p[3] = {x, y, z};
  const int e[3] = {b.nx, b.ny, b.nz};
  p[ax] += o;
  if (p[ax] < 0 || p[ax] >= e[ax]) {
    if (!b.per[ax]) return -1;
    p[ax] = (p[ax] % e[ax] + e[ax]) % e[ax];
  }
  return (long long)p[0] + (long long)e[0] * ((long long)p[1] + (long long)e[1] * p[2]);
}

// stencil nodes along axis ax, (x, x+1, x+2, x-1, x-2), and which are usable
__device__ void stencil(const Box& b, const int32_t* __restrict__ state,
                        const uint8_t* __restrict__ filled, long long k, int x, int y, int z,
                        int ax, long long nd[5], bool ok[5]) {
  nd[0] = k;
  ok[0] = true;
  const int off[4] = {1, 2, -1, -2};
  for (int m = 0; m < 4; ++m) {
    nd[m + 1] = nb_node(b, x, y, z, ax, off[m]);
    ok[m + 1] = nd[m + 1] >= 0;
  }
  if (state) {
    const int dir[4] = {1 + ax, 15 + ax, 4 + ax, 18 + ax};
    for (int m = 0; m < 4; ++m) ok[m + 1] = ok[m + 1] && state[(size_t)dir[m] * b.n + k] >= 0;
  }
  if (filled)
    for (int m = 1; m < 5; ++m) ok[m] = ok[m] && filled[nd[m]] != 0;
}

template <bool kShifted>
__global__ __launch_bounds__(256) void k27_tag_fields(const void* __restrict__ f_, Box b,
                                                      int quantity, double* __restrict__ q,
                                                      double* __restrict__ vel) {
  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= b.n) return;
  double r = 0.0, j[3] = {0.0, 0.0, 0.0};
  for (int i = 0; i < 27; ++i) {
    double v;
    if constexpr (kShifted) v = (double)static_cast<const float*>(f_)[(size_t)i * b.n + k];
    else v = static_cast<const double*>(f_)[(size_t)i * b.n + k];
    r += v;
    j[0] += CX27[i] * v;
    j[1] += CY27[i] * v;
    j[2] += CZ27[i] * v;
  }
  if constexpr (kShifted) r += 1.0;
  const double u = j[0] / r, v = j[1] / r, w = j[2] / r;
  vel[k] = u;
  vel[b.n + k] = v;
  vel[2 * b.n + k] = w;
  q[k] = quantity == 1 ? r
         : quantity == 2 ? u
         : quantity == 3 ? v
         : quantity == 4 ? w
                         : sqrt(u * u + v * v + w * w);
}

// |grad src| per node; 0 on nodes whose code is < 0 (skip_null: except -666,
// as calculate_grad_mag.f95:37 has it)
// (filled, when given: only filled positions are computed or usable)
__global__ __launch_bounds__(256) void k27_tag_gradmag(const double* __restrict__ src,
                                                       const double* __restrict__ vel, Box b,
                                                       const int32_t* __restrict__ state,
                                                       const uint8_t* __restrict__ filled,
                                                       int scheme, int skip_null,
                                                       double* __restrict__ out) {
  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= b.n) return;
  if ((filled && !filled[k]) || (state && state[k] < 0 && !(skip_null && state[k] == -666))) {
    out[k] = 0.0;
    return;
  }
  const int x = (int)(k % b.nx), y = (int)((k / b.nx) % b.ny), z = (int)(k / ((long long)b.nx * b.ny));
  double d[3];
  for (int ax = 0; ax < 3; ++ax) {
    long long nd[5];
    bool ok[5];
    stencil(b, state, filled, k, x, y, z, ax, nd, ok);
    double qv[5];
    for (int m = 0; m < 5; ++m) qv[m] = nd[m] >= 0 ? src[nd[m]] : 0.0;
    d[ax] = windage2(qv, ok, vel[(size_t)ax * b.n + k], scheme);
  }
  out[k] = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
}

// tags of the valid nodes (local box positions vlo..vhi), 0 elsewhere
__global__ __launch_bounds__(256) void k27_tag_mark(const double* __restrict__ ggm, Box b,
                                                    double limit, int3 vlo, int3 vhi,
                                                    uint8_t* __restrict__ tags) {
  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= b.n) return;
  const int x = (int)(k % b.nx), y = (int)((k / b.nx) % b.ny), z = (int)(k / ((long long)b.nx * b.ny));
  const bool valid = x >= vlo.x && x <= vhi.x && y >= vlo.y && y <= vhi.y && z >= vlo.z && z <= vhi.z;
  tags[k] = valid && ggm[k] >= limit ? 1 : 0;
}

// A level of many boxes on a dense array over the bounding box of its stored
// nodes: pass 0 places the ghost-shell nodes (filled 1, every link usable),
// pass 1 the valid nodes over them (filled 2, the level's link codes).  A
// position no box stores stays unfilled: never computed, never usable.
__global__ __launch_bounds__(256) void k27_tag_scatter(
    const double* __restrict__ q_s, const double* __restrict__ vel_s, long long ns,
    const int32_t* __restrict__ state_s, const BoxDesc* __restrict__ boxes,
    const int* __restrict__ blo, int3 dlo, Box d, int pass, double* __restrict__ q_d,
    double* __restrict__ vel_d, uint8_t* __restrict__ filled, int32_t* __restrict__ state_d) {
  const unsigned box = blockIdx.y;
  const BoxDesc bx = boxes[box];
  const unsigned t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= bx.n) return;
  const unsigned nz = bx.n / (bx.nx * bx.ny);
  const unsigned lx = t % bx.nx, ly = (t / bx.nx) % bx.ny, lz = t / (bx.nx * bx.ny);
  const bool shell = lx == 0 || ly == 0 || lz == 0 || lx == bx.nx - 1 || ly == bx.ny - 1 ||
                     lz == nz - 1;
  if (shell != (pass == 0)) return;
  const long long gx = blo[3 * box] - 1 + (int)lx - dlo.x, gy = blo[3 * box + 1] - 1 + (int)ly - dlo.y,
                  gz = blo[3 * box + 2] - 1 + (int)lz - dlo.z;
  const long long kd = gx + (long long)d.nx * (gy + (long long)d.ny * gz);
  const long long ks = (long long)bx.off + t;
  q_d[kd] = q_s[ks];
  for (int c = 0; c < 3; ++c) vel_d[c * d.n + kd] = vel_s[c * ns + ks];
  filled[kd] = pass == 0 ? 1 : 2;
  if (state_d)
    for (int i = 0; i < 27; ++i)
      state_d[(size_t)i * d.n + kd] = pass == 0 || !state_s ? 0 : state_s[(size_t)i * ns + ks];
}

// dense ggm and tags back to the stored nodes (valid nodes; 0 on the shell)
__global__ __launch_bounds__(256) void k27_tag_gather(
    const double* __restrict__ ggm_d, const uint8_t* __restrict__ tag_d,
    const BoxDesc* __restrict__ boxes, const int* __restrict__ blo, int3 dlo, Box d,
    double* __restrict__ ggm_s, uint8_t* __restrict__ tag_s) {
  const unsigned box = blockIdx.y;
  const BoxDesc bx = boxes[box];
  const unsigned t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= bx.n) return;
  const unsigned nz = bx.n / (bx.nx * bx.ny);
  const unsigned lx = t % bx.nx, ly = (t / bx.nx) % bx.ny, lz = t / (bx.nx * bx.ny);
  const bool shell = lx == 0 || ly == 0 || lz == 0 || lx == bx.nx - 1 || ly == bx.ny - 1 ||
                     lz == nz - 1;
  const long long ks = (long long)bx.off + t;
  if (shell) {
    ggm_s[ks] = 0.0;
    tag_s[ks] = 0;
    return;
  }
  const long long gx = blo[3 * box] - 1 + (int)lx - dlo.x, gy = blo[3 * box + 1] - 1 + (int)ly - dlo.y,
                  gz = blo[3 * box + 2] - 1 + (int)lz - dlo.z;
  const long long kd = gx + (long long)d.nx * (gy + (long long)d.ny * gz);
  ggm_s[ks] = ggm_d[kd];
  tag_s[ks] = tag_d[kd];
}

// tags of the filled-2 (valid) positions of the dense array
__global__ __launch_bounds__(256) void k27_tag_mark_dense(const double* __restrict__ ggm,
                                                          const uint8_t* __restrict__ filled,
                                                          long long n, double limit,
                                                          uint8_t* __restrict__ tags) {
  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  tags[k] = filled[k] == 2 && ggm[k] >= limit ? 1 : 0;
}

unsigned tag_blocks(long long n) { return (unsigned)((n + 255) / 256); }

}  // namespace

// work: 5 n doubles (q, velocity, gm); ggm: n doubles; tags: n bytes
hipError_t launch_d3q27_ggm(const Box& b, const void* f, int shifted, const int32_t* state,
                            int quantity, int scheme, double limit, const int vlo[3],
                            const int vhi[3], double* work, double* ggm, uint8_t* tags,
                            hipStream_t s) {
  double* q = work;
  double* vel = work + b.n;
  double* gm = work + 4 * b.n;
  const unsigned g = tag_blocks(b.n);
  if (shifted) k27_tag_fields<true><<<g, 256, 0, s>>>(f, b, quantity, q, vel);
  else k27_tag_fields<false><<<g, 256, 0, s>>>(f, b, quantity, q, vel);
  k27_tag_gradmag<<<g, 256, 0, s>>>(q, vel, b, state, nullptr, scheme, 1, gm);
  k27_tag_gradmag<<<g, 256, 0, s>>>(gm, vel, b, state, nullptr, scheme, 0, ggm);
  k27_tag_mark<<<g, 256, 0, s>>>(ggm, b, limit, make_int3(vlo[0], vlo[1], vlo[2]),
                                 make_int3(vhi[0], vhi[1], vhi[2]), tags);
  return hipGetLastError();
}

// A refined level of many boxes: the fields on the stored nodes (stored work
// 4 ns doubles), placed on the dense bounding box d (origin dlo in level
// indices; dense work 6 nd doubles, filled and tags nd bytes each, state_d
// 27 nd int32 when the level has a link table), the indicator there, back to
// the stored nodes.  blo: [nbox][3] device copy of the boxes' valid lo.
hipError_t launch_d3q27_ggm_boxes(const void* f, long long ns, int shifted, const int32_t* state,
                                  const BoxDesc* boxes, const int* blo, unsigned nbox,
                                  unsigned max_box_n, const int dlo[3], const Box& d, int quantity,
                                  int scheme, double limit, double* work_s, double* work_d,
                                  uint8_t* filled, uint8_t* tag_d, int32_t* state_d,
                                  double* ggm_s, uint8_t* tag_s, hipStream_t s) {
  Box bs{};
  bs.nx = 1;
  bs.ny = 1;
  bs.nz = 1;
  bs.n = ns;
  double* q_s = work_s;
  double* vel_s = work_s + ns;
  double* q_d = work_d;
  double* vel_d = work_d + d.n;
  double* gm_d = work_d + 4 * d.n;
  double* ggm_d = work_d + 5 * d.n;
  const unsigned gs = tag_blocks(ns), gd = tag_blocks(d.n);
  if (shifted) k27_tag_fields<true><<<gs, 256, 0, s>>>(f, bs, quantity, q_s, vel_s);
  else k27_tag_fields<false><<<gs, 256, 0, s>>>(f, bs, quantity, q_s, vel_s);
  hipError_t e = launch_fill(work_d, 4 * d.n, 0.0, s);
  if (e == hipSuccess) e = hipMemsetAsync(filled, 0, (size_t)d.n, s);
  if (e != hipSuccess) return e;
  const dim3 grid(tag_blocks(max_box_n), nbox);
  const int3 o = make_int3(dlo[0], dlo[1], dlo[2]);
  for (int pass = 0; pass < 2; ++pass)
    k27_tag_scatter<<<grid, 256, 0, s>>>(q_s, vel_s, ns, state, boxes, blo, o, d, pass, q_d, vel_d,
                                         filled, state_d);
  k27_tag_gradmag<<<gd, 256, 0, s>>>(q_d, vel_d, d, state_d, filled, scheme, 1, gm_d);
  k27_tag_gradmag<<<gd, 256, 0, s>>>(gm_d, vel_d, d, state_d, filled, scheme, 0, ggm_d);
  k27_tag_mark_dense<<<gd, 256, 0, s>>>(ggm_d, filled, d.n, limit, tag_d);
  k27_tag_gather<<<grid, 256, 0, s>>>(ggm_d, tag_d, boxes, blo, o, d, ggm_s, tag_s);
  return hipGetLastError();
}

}  // namespace uclbs
