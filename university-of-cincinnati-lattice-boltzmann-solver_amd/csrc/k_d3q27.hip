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

#include <algorithm>

namespace uclbs {

namespace {

__device__ __forceinline__ void wave_stats(unsigned long long* __restrict__ stats, bool active,
                                           int iters, int status, bool fallback,
                                           bool logpath) {
  const unsigned long long act = __ballot(active);
  if (act == 0ull) return;
  // status -1: the node is finished elsewhere; only its iterations count here
  const unsigned long long done = __ballot(active && status >= 0);
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
    if (done) atomicAdd(s + ST_NODES, (unsigned long long)__popcll(done));
    if (it) atomicAdd(s + ST_ITERS, it);
    if (nt) atomicAdd(s + ST_TRIVIAL, nt);
    if (nb) atomicAdd(s + ST_BOUND, nb);
    if (nn) atomicAdd(s + ST_NONPOS, nn);
    if (nc) atomicAdd(s + ST_NOCONV, nc);
    if (nf) atomicAdd(s + ST_FALLBACK, nf);
    if (nl) atomicAdd(s + ST_LOGPATH, nl);
  }
}

// f[i * nn + k] through a wave-uniform population base and a 32-bit byte
// offset (FP32: boxes hold under 2^30 nodes), which the compiler issues as one
// global load / store with a scalar base and no per-lane 64-bit address
template <typename T>
__device__ __forceinline__ const T& pop(const T* __restrict__ f, unsigned i, unsigned nn,
                                        unsigned k) {
  if constexpr (sizeof(T) == 4) {
    const char* base = reinterpret_cast<const char*>(f + (size_t)i * nn);
    return *reinterpret_cast<const T*>(base + k * 4u);
  } else {
    return f[(size_t)i * nn + k];
  }
}
template <typename T>
__device__ __forceinline__ T& pop(T* __restrict__ f, unsigned i, unsigned nn, unsigned k) {
  return const_cast<T&>(pop(const_cast<const T*>(f), i, nn, k));
}
// the step's output populations: a streaming store (nontemporal), since no
// kernel of this step reads them back and each line is written once
template <typename T>
__device__ __forceinline__ void pop_stream(T* __restrict__ f, unsigned i, unsigned nn, unsigned k,
                                           T v) {
  __builtin_nontemporal_store(v, &pop(f, i, nn, k));
}

// appends the wavefront's deferred nodes, with their alpha seeds, to shard
// `shard` of the far-from-equilibrium queue (one atomic per wavefront; the
// shards keep the streaming kernel's atomics off a single address)
__device__ __forceinline__ void queue_far(const StepArgs& a, unsigned shard, bool defer,
                                          unsigned node, double seed) {
  const unsigned long long dm = __ballot(defer);
  if (dm) {
    unsigned base = 0;
    const int leader = __ffsll((long long)dm) - 1;
    if ((int)__lane_id() == leader) base = atomicAdd(a.defer_count + shard, (unsigned)__popcll(dm));
    base = __shfl(base, leader);
    if (defer)
      a.defer_list[(size_t)shard * a.defer_cap + base + __popcll(dm & ((1ull << __lane_id()) - 1ull))] =
          make_uint2(node, __float_as_uint((float)seed));
  }
}

// kFlags: the level has link tables; only interior fluid nodes (flag 0) are
// updated here, the boundary kernel updates the rest.
template <bool kShifted, bool kKeepAlpha, bool kFlags>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k27_step_periodic(const void* __restrict__ fin_,
                                                         void* __restrict__ fout_,
                                                         double* __restrict__ alpha_out,
                                                         unsigned long long* __restrict__ stats,
                                                         Box b, StepArgs a,
                                                         const uint8_t* __restrict__ flags,
                                                         BcTables bcf) {
  using T = typename std::conditional<kShifted, float, double>::type;
  const T* __restrict__ fin = static_cast<const T*>(fin_);
  T* __restrict__ fout = static_cast<T*>(fout_);
  // 32-bit node offsets against a per-population scalar base (boxes are
  // limited to 2^30 nodes), so each of the 27 loads needs one address VGPR
  const unsigned gid = blockIdx.x * blockDim.x + threadIdx.x;
  const unsigned nn = (unsigned)b.n;
  const unsigned node = gid < nn ? gid : 0u;
  bool active = gid < nn;
  uint8_t fl = 0;
  if constexpr (kFlags) {
    fl = flags[node];
    active = active && (fl == kFlagInterior || fl == kFlagFace);
  }
  const unsigned nx = (unsigned)b.nx, ny = (unsigned)b.ny, nz = (unsigned)b.nz;
  const unsigned x = node % nx;
  const unsigned yz = node / nx;
  const unsigned y = yz % ny;
  const unsigned z = yz / ny;
  const unsigned sy = nx, sz = nx * ny;
  // source coordinate x - c for c = -1, 0, +1 (periodic wrap)
  unsigned ox[3], oy[3], oz[3];
  ox[0] = x + 1 == nx ? 0u : x + 1;
  ox[1] = x;
  ox[2] = x == 0 ? nx - 1 : x - 1;
  oy[0] = (y + 1 == ny ? 0u : y + 1) * sy;
  oy[1] = y * sy;
  oy[2] = (y == 0 ? ny - 1 : y - 1) * sy;
  oz[0] = (z + 1 == nz ? 0u : z + 1) * sz;
  oz[1] = z * sz;
  oz[2] = (z == 0 ? nz - 1 : z - 1) * sz;

  T f[27];
#pragma unroll
  for (int i = 0; i < 27; ++i) f[i] = pop(fin, i, nn, ox[CX27[i] + 1] + oy[CY27[i] + 1] + oz[CZ27[i] + 1]);
  // Nodes whose only boundary links leave the box through a face: the
  // population arriving from outside takes the face's code, the same code the
  // link table holds (first non-periodic face in the order x-, y-, z-, x+, y+,
  // z+, so faces are applied in reverse order and the first one wins).
  // -10: halfway bounce-back of the node's own opposite population; inlet /
  // outlet / freestream: their equilibrium.  Wave-uniform guards keep the
  // interior wavefronts out of this code.
  if constexpr (kFlags) {
    if (__any(fl == kFlagFace)) {
      const bool isf = fl == kFlagFace;
      const bool on[6] = {isf && x == 0u && !b.per[0], isf && y == 0u && !b.per[1],
                          isf && z == 0u && !b.per[2], isf && x == nx - 1 && !b.per[0],
                          isf && y == ny - 1 && !b.per[1], isf && z == nz - 1 && !b.per[2]};
#pragma unroll
      for (int q = 5; q >= 0; --q) {
        if (!__any(on[q])) continue;
        const int32_t code = a.face_code[q];
        const double* tab = code == -10 ? nullptr
                            : (code <= -101 && code >= -199) ? bcf.feq_in + (-code - 101) * 27
                            : (code <= -201 && code >= -299) ? bcf.feq_out + (-code - 201) * 27
                                                             : bcf.feq_fs;
#pragma unroll
        for (int i = 1; i < 27; ++i) {
          // the source x - c_i lies beyond face q
          const int cq = q % 3 == 0 ? CX27[i] : (q % 3 == 1 ? CY27[i] : CZ27[i]);
          if (!(q < 3 ? cq > 0 : cq < 0)) continue;
          T v;
          if (tab) {
            if constexpr (kShifted) v = (T)(tab[i] - W27[i]);
            else v = (T)tab[i];
          } else {
            v = on[q] ? pop(fin, OPP27[i], nn, node) : f[i];
          }
          if (on[q]) f[i] = v;
        }
      }
    }
  }

  AlphaCfg ac{a.tol, a.maxit, a.alpha_mode, a.thr};
  double rho, u[3];
  AlphaOut ao;
  // learned mode: the initial guess left in alpha[] by the guess kernel
  const double a0 = (kShifted && a.alpha_mode == 2 && active) ? alpha_out[node] : 2.0;
  if constexpr (kShifted) ao = collide27_shifted(f, a.beta, ac, a0, rho, u);
  else ao = collide27_fp64(f, a.beta, ac, 2.0, rho, u);

  if (active) {
    // a deferred node stores its pre-collision state; the far-from-equilibrium
    // kernel collides it in place
#pragma unroll
    for (int i = 0; i < 27; ++i) pop_stream(fout, i, nn, node, f[i]);
    if (kKeepAlpha) alpha_out[node] = ao.alpha;
  }
  if constexpr (kShifted)
    queue_far(a, blockIdx.x % kFarShards, active && ao.status == ALPHA_DEFER, node, ao.alpha);
  // deferred nodes report their seed iterations here and the rest in the far kernel
  wave_stats(stats, active, ao.iters, ao.status == ALPHA_DEFER ? -1 : ao.status, ao.fallback,
             ao.logpath);
}

// All boxes of a level in one launch: block b takes tile b (tile_nodes nodes
// of one box), pulls over the box's strides -- a valid node's neighbours lie in
// the box or its ghost shell -- and collides as k27_step_periodic does.  A
// ghost node that shows a valid node s of the level (kFlagMirror) repeats s's
// update -- the pull around s in s's box and the same collide, so the same
// bits -- and stores it in place: the shell is current after the step with no
// copy pass (a deferred mirror is queued too, marked, and solved by the far
// kernel like s).  Ghosts at a coarse-fine interface are interpolated before
// the step.
template <bool kShifted, bool kKeepAlpha, bool kLinks>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k27_step_boxes(
    const void* __restrict__ fin_, void* __restrict__ fout_, double* __restrict__ alpha_out,
    unsigned long long* __restrict__ stats, unsigned nn, StepArgs a,
    const uint8_t* __restrict__ flags, const uint4* __restrict__ alias,
    const BoxTile* __restrict__ tiles, const BoxDesc* __restrict__ boxes, int tile_nodes) {
  using T = typename std::conditional<kShifted, float, double>::type;
  const T* __restrict__ fin = static_cast<const T*>(fin_);
  T* __restrict__ fout = static_cast<T*>(fout_);
  const BoxTile t = tiles[blockIdx.x];
  const BoxDesc bx = boxes[t.box];
  const unsigned local = t.start + threadIdx.x;
  const bool in_box = (int)threadIdx.x < tile_nodes && local < bx.n;
  const unsigned node = bx.off + (in_box ? local : 0u);
  // the valid nodes are the inside of the stored box, the ghost shell its
  // outer layer: known from the coordinates, so only shell lanes read their
  // flag and mirror entry (together), and the pulls of a wavefront wait for
  // one load instead of a chain of three
  const unsigned sxy = bx.nx * bx.ny;
  const unsigned lx = local % bx.nx, lyz = local / bx.nx, ly = lyz % bx.ny, lz = lyz / bx.ny;
  const bool shell = lx == 0u || ly == 0u || lz == 0u || lx + 1u == bx.nx || ly + 1u == bx.ny ||
                     lz + 1u == bx.n / sxy;
  uint8_t fl = kFlagInterior;
  uint4 al = make_uint4(0u, 0u, 0u, 0u);
  if (in_box && shell) {
    fl = flags[node];
    al = alias[node];
  } else if (kLinks && in_box) {
    // with a link table a valid node may be a boundary node (stepped by the
    // boundary kernel) or solid (left as it is)
    fl = flags[node];
  }
  const bool active = in_box && !shell && fl == kFlagInterior;  // counted node updates
  const bool mirror = in_box && shell && fl == kFlagMirror;
  // the node whose update this lane computes, and that node's box strides
  const unsigned src = mirror ? al.x : node;
  const unsigned sy = mirror ? al.y : bx.nx;
  const unsigned sz = mirror ? al.z : sxy;
  const bool pull = active || mirror;
  T f[27];
#pragma unroll
  for (int i = 0; i < 27; ++i)
    f[i] = pop(fin, i, nn, pull ? src - CX27[i] - CY27[i] * sy - CZ27[i] * sz : node);
  AlphaCfg ac{a.tol, a.maxit, a.alpha_mode, a.thr};
  double rho, u[3];
  AlphaOut ao;
  if constexpr (kShifted) ao = collide27_shifted(f, a.beta, ac, 2.0, rho, u);
  else ao = collide27_fp64(f, a.beta, ac, 2.0, rho, u);
  // ghost-shell lanes store too (values nobody reads: the shell is refilled
  // before the next step), so every line of the box is written whole and the
  // L2 does not fetch partly written lines; boundary and solid valid nodes
  // are written by the boundary kernel or not at all
  if (in_box && (!kLinks || shell || active)) {
#pragma unroll
    for (int i = 0; i < 27; ++i) pop_stream(fout, i, nn, node, f[i]);
  }
  if (kKeepAlpha && active) alpha_out[node] = ao.alpha;
  if constexpr (kShifted)
    queue_far(a, blockIdx.x % kFarShards, pull && ao.status == ALPHA_DEFER,
              mirror ? (node | kMirrorBit) : node, ao.alpha);
  wave_stats(stats, active, ao.iters, ao.status == ALPHA_DEFER ? -1 : ao.status, ao.fallback,
             ao.logpath);
}

// Far-from-equilibrium nodes queued by the streaming and boundary kernels
// (max|x| beyond the in-register series): one thread per queued node reads
// the pre-collision state stored in place, re-forms D, and solves for alpha by
// the re-centred logarithmic sweeps from the series seed left in alpha[]; then
// relaxes and stores.  The queue length stays on the device; the grid is
// capped and strides over the queue, and blocks past its end leave at once.
// Block 0 zeroes the other set of lengths, which the next step appends to.
// One-wave blocks: the 13.8 KB of LDS per block lets three waves share a SIMD.
__global__ __launch_bounds__(kFarBlock) void k27_far(float* __restrict__ f_, double* __restrict__ alpha,
                                               unsigned long long* __restrict__ stats, Box b,
                                               StepArgs a, int keep_alpha) {
  // the next step's lengths start from zero (the previous far kernel read them)
  if (blockIdx.x == 0)
    for (unsigned s = threadIdx.x; s <= (unsigned)kFarShards; s += blockDim.x) a.defer_next[s] = 0u;
  // the logarithm table, one entry per lane (kFarBlock = 64 entries)
  __shared__ LogTable lt;
  log_table_entry(lt, (int)threadIdx.x);
  // exclusive prefix of the shard lengths: lane l sums shards 5l .. 5l+4
  __shared__ unsigned pre[kFarShards + 2];
  {
    constexpr int kPer = (kFarShards + 1 + 63) / 64;
    static_assert(kFarBlock == 64 && kPer * 64 >= kFarShards + 1, "one wave scans the shards");
    const int l = (int)threadIdx.x;
    unsigned c[kPer], sum = 0;
#pragma unroll
    for (int q = 0; q < kPer; ++q) {
      const int s = l * kPer + q;
      c[q] = s <= kFarShards ? a.defer_count[s] : 0u;
      sum += c[q];
    }
    unsigned inc = sum;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const unsigned v = __shfl_up(inc, d);
      if (l >= d) inc += v;
    }
    unsigned run = inc - sum;
#pragma unroll
    for (int q = 0; q < kPer; ++q) {
      const int s = l * kPer + q;
      if (s <= kFarShards + 1) pre[s] = run;
      run += c[q];
    }
    __syncthreads();
  }
  const unsigned count = pre[kFarShards + 1];
  const unsigned stride = gridDim.x * blockDim.x;
  const unsigned first = blockIdx.x * blockDim.x;
  const unsigned nn = (unsigned)b.n;
  AlphaCfg ac{a.tol, a.maxit, a.alpha_mode, a.thr};
  __shared__ float sf[27 * kFarBlock], sD[27 * kFarBlock];
  // grid-stride over the concatenated shards; whole blocks past the end leave at once
  for (unsigned base = first; base < count; base += stride) {
    const unsigned k = base + threadIdx.x;
    const bool active = k < count;
    const unsigned kk = active ? k : base;
    // shard holding entry kk: the last s with pre[s] <= kk
    int lo = 0, hi = kFarShards;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (pre[mid] <= kk) lo = mid; else hi = mid - 1;
    }
    const uint2 e = a.defer_list[(size_t)lo * a.defer_cap + (kk - pre[lo])];
    const unsigned node = e.x & ~kMirrorBit;
    const bool counted = active && !(e.x & kMirrorBit);
    float f[27];
#pragma unroll
    for (int i = 0; i < 27; ++i) f[i] = (f_ + (size_t)i * nn)[node];
    const double seed = (double)__uint_as_float(e.y);
    AlphaOut ao = collide27_far(f, a.beta, ac, seed, sf + threadIdx.x, sD + threadIdx.x, active,
                                f_, nn, node, lt);
    if (counted && keep_alpha) alpha[node] = ao.alpha;
    wave_stats(stats, counted, ao.iters, ao.status, false, ao.logpath);
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

// density and momentum of stored node k of a level of n nodes
template <bool kShifted>
__device__ __forceinline__ void moments27(const void* __restrict__ src_, int64_t n, int64_t k,
                                          double& r, double& jx, double& jy, double& jz) {
  r = jx = jy = jz = 0.0;
#pragma unroll
  for (int i = 0; i < 27; ++i) {
    double v;
    if constexpr (kShifted) v = (double)static_cast<const float*>(src_)[(int64_t)i * n + k];
    else v = static_cast<const double*>(src_)[(int64_t)i * n + k];
    r += v;
    jx += CX27[i] * v;
    jy += CY27[i] * v;
    jz += CZ27[i] * v;
  }
  if constexpr (kShifted) r += 1.0;
}

template <bool kShifted>
__global__ void k27_macro(const void* __restrict__ src_, Box b, double* __restrict__ rho,
                          double* __restrict__ ux, double* __restrict__ uy,
                          double* __restrict__ uz) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= b.n) return;
  double r, jx, jy, jz;
  moments27<kShifted>(src_, b.n, k, r, jx, jy, jz);
  rho[k] = r;
  ux[k] = jx / r;
  uy[k] = jy / r;
  uz[k] = jz / r;
}

// flow variables at listed stored nodes: out[6][m] = rho, u, v, w, T, p with
// the isothermal lattice temperature T = 1/3 and p = rho (T + |u|^2 / 2), the
// pressure column of calculate_flow_variables.f95:220-221
template <bool kShifted>
__global__ void k27_probe(const void* __restrict__ src_, int64_t n,
                          const uint32_t* __restrict__ nodes, unsigned m,
                          double* __restrict__ out) {
  const unsigned j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= m) return;
  double r, jx, jy, jz;
  moments27<kShifted>(src_, n, (int64_t)nodes[j], r, jx, jy, jz);
  const double u = jx / r, v = jy / r, w = jz / r, T = 1.0 / 3.0;
  out[j] = r;
  out[m + j] = u;
  out[2 * m + j] = v;
  out[3 * m + j] = w;
  out[4 * m + j] = T;
  out[5 * m + j] = r * (T + (u * u + v * v + w * w) / 2.0);
}

// Taylor-Green vortex: u = U0 (sin X cos Y cos Z, -cos X sin Y cos Z, 0) plus
// uniform noise in (-amp, amp) per component from Philox4x32-10 keyed by the
// seed and counted by (node, component); f = feq(1, u).
// Region form: the stored nodes of a box of extents ext whose node (0,0,0)
// sits at domain position org (wrapped into the periodic domain dom), written
// from node out_off of arrays of nn nodes; the field and the noise counter
// follow the domain position, so a domain split into boxes starts from
// exactly the state of the domain as one box.
template <bool kShifted>
__global__ void k27_init_tgv(void* __restrict__ dst_, Box dom, int64_t nn, int64_t out_off,
                             int ex, int ey, int ez, int ox, int oy, int oz, double u0, double amp,
                             uint64_t seed) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= (int64_t)ex * ey * ez) return;
  const int lx = (int)(r % ex);
  const int64_t lyz = r / ex;
  const int ly = (int)(lyz % ey), lz = (int)(lyz / ey);
  const int x = ((ox + lx) % dom.nx + dom.nx) % dom.nx, y = ((oy + ly) % dom.ny + dom.ny) % dom.ny,
            z = ((oz + lz) % dom.nz + dom.nz) % dom.nz;
  const int64_t kd = x + (int64_t)dom.nx * (y + (int64_t)dom.ny * z);
  const int64_t k = out_off + r;
  const double tp = 6.283185307179586476925286766559;
  const double X = tp * x / dom.nx, Y = tp * y / dom.ny, Z = tp * z / dom.nz;
  double u[3];
  u[0] = u0 * sin(X) * cos(Y) * cos(Z);
  u[1] = -u0 * cos(X) * sin(Y) * cos(Z);
  u[2] = 0.0;
  for (int c = 0; c < 3; ++c) u[c] += amp * (2.0 * philox_uniform((uint64_t)kd, (uint32_t)c, seed) - 1.0);
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
    if constexpr (kShifted) static_cast<float*>(dst_)[(int64_t)i * nn + k] = (float)(feq - W27[i]);
    else static_cast<double*>(dst_)[(int64_t)i * nn + k] = feq;
  }
}


// ------------------------------------------------------------------ boundaries
// One thread per fluid node that has a boundary link.  The incoming population
// f_i(x) (velocity c_i, from x - c_i) is chosen by the code s of the link from
// x towards x - c_i, state[opp(i)][x] -- the dispatch of streaming.f95:365-1133
// seen from the receiving node:
//   s >= 0         pull fout_i(x - c_i)
//   s = -10        halfway bounce-back fout_opp(i)(x) (the rule bounceback.f95
//                  aims at; its in-place swap is not reproduced)
//   -1xx / -2xx    inlet / outlet equilibrium table (inlet_bc.f95, outlet_bc.f95,
//                  rho_v_bc.f95: fi(x,opp(k)) = inl_fieq(opp(k)))
//   -1000          freestream equilibrium (freestream_edge_bc.f95:18-19)
//   <= -1001       linear interpolated bounce-back (Bouzidi, Firdaouss &
//                  Lallemand 2001) on the post-collision populations: the wall
//                  lies at q = |s + 1001|/1e9 (target_finder.f95:40) of the link
//                  from x towards x - c_i, k = opp(i) points into it;
//                  q < 1/2: f_i = 2q fout_k(x) + (1 - 2q) fout_k(x + c_i),
//                  q >= 1/2: f_i = fout_k(x)/(2q) + (1 - 1/(2q)) fout_i(x),
//                  halfway bounce-back where x + c_i is not a fluid node
//                  (the reference's D3Q27 lattice has no wall rule; both forms
//                  are convex combinations, so no population turns negative)
//   other          fout_opp(i)(x)
// The node then collides with the streaming kernel's collide.  All BC
// arithmetic is FP64.
template <bool kShifted>
__device__ __forceinline__ double ldf(const void* f, int i, int64_t n, int64_t k) {
  if constexpr (kShifted) return W27[i] + (double)static_cast<const float*>(f)[(int64_t)i * n + k];
  else return static_cast<const double*>(f)[(int64_t)i * n + k];
}

__device__ __forceinline__ int wrapc(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }

__device__ void feq27_fp64(double rho, const double u[3], double* feq) {
  double p[3][3];
  for (int d = 0; d < 3; ++d) {
    const double s = sqrt(1.0 + 3.0 * u[d] * u[d]);
    const double bb = 2.0 - s;
    p[d][1] = bb;
    p[d][2] = bb * (2.0 * u[d] + s) / (1.0 - u[d]);
    p[d][0] = bb * (1.0 - u[d]) / (2.0 * u[d] + s);
  }
#pragma unroll
  for (int i = 0; i < 27; ++i)
    feq[i] = rho * W27[i] * p[0][CX27[i] + 1] * p[1][CY27[i] + 1] * p[2][CZ27[i] + 1];
}

template <bool kShifted>
__global__ __launch_bounds__(64) void k27_boundary(const void* __restrict__ fin,
                                                    void* __restrict__ fout_,
                                                    double* __restrict__ alpha, int keep_alpha,
                                                    unsigned long long* __restrict__ stats,
                                                    Box b, StepArgs a, BcArgs bc) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool active = j < bc.nlist;
  const int64_t k = bc.list[active ? j : 0];
  const int64_t n = b.n;
  // the node's box: the level's one box (periodic axes wrap) or, on a level
  // of many boxes, the box whose ghost shell holds the node's neighbours
  int64_t base = 0;
  int bnx = b.nx, bny = b.ny, bnz = b.nz;
  if (bc.boxes) {
    const BoxDesc bx = bc.boxes[bc.box_of[active ? j : 0]];
    base = bx.off;
    bnx = (int)bx.nx;
    bny = (int)bx.ny;
    bnz = (int)(bx.n / (bx.nx * bx.ny));
  }
  const int64_t lk = k - base;
  const int x = (int)(lk % bnx);
  const int64_t yz = lk / bnx;
  const int y = (int)(yz % bny), z = (int)(yz / bny);
  double f[27];
  bool unknown[27];
  bool any_wall = false;
#pragma unroll
  for (int i = 0; i < 27; ++i) {
    unknown[i] = false;
    if (i == 0) { f[0] = ldf<kShifted>(fin, 0, n, k); continue; }
    const int32_t s = bc.state[(int64_t)OPP27[i] * n + k];
    if (s >= 0) {
      const int64_t src = base + wrapc(x - CX27[i], bnx) +
                          (int64_t)bnx * (wrapc(y - CY27[i], bny) + (int64_t)bny * wrapc(z - CZ27[i], bnz));
      f[i] = ldf<kShifted>(fin, i, n, src);
    } else if (s == -10) {
      f[i] = ldf<kShifted>(fin, OPP27[i], n, k);
    } else if (s <= -100 && s >= -199 && -s - 100 <= bc.n_in) {
      f[i] = bc.feq_in[(int64_t)(-s - 101) * 27 + i];
    } else if (s <= -200 && s >= -299 && -s - 200 <= bc.n_out) {
      f[i] = bc.feq_out[(int64_t)(-s - 201) * 27 + i];
    } else if (s == -1000 && bc.feq_fs) {
      f[i] = bc.feq_fs[i];
    } else if (s <= -1001) {
      unknown[i] = true;
      any_wall = true;
      f[i] = 0.0;
    } else {
      f[i] = ldf<kShifted>(fin, OPP27[i], n, k);
    }
  }
  int nibb = 0;
  if (any_wall && active) {
#pragma unroll 1
    for (int i = 1; i < 27; ++i) {
      if (!unknown[i]) continue;
      const int kk = OPP27[i];
      const double q = fabs((double)bc.state[(int64_t)kk * n + k] + 1001.0) / 1.0e9;
      const double fk = ldf<kShifted>(fin, kk, n, k);
      double v;
      if (q >= 0.5) {
        v = fk / (2.0 * q) + (1.0 - 1.0 / (2.0 * q)) * ldf<kShifted>(fin, i, n, k);
      } else if (bc.state[(int64_t)i * n + k] >= 0) {
        // the fluid node behind x, x + c_i
        const int64_t src = base + wrapc(x + CX27[i], bnx) +
                            (int64_t)bnx * (wrapc(y + CY27[i], bny) + (int64_t)bny * wrapc(z + CZ27[i], bnz));
        v = 2.0 * q * fk + (1.0 - 2.0 * q) * ldf<kShifted>(fin, kk, n, src);
      } else {
        v = fk;
      }
      f[i] = v;
      nibb++;
    }
  }
  AlphaCfg ac{a.tol, a.maxit, a.alpha_mode, a.thr};
  double rho, u[3];
  AlphaOut ao;
  if constexpr (kShifted) {
    // the streaming kernel's collide: a node beyond the series tiers stores
    // its pre-collision state and joins the far-from-equilibrium queue
    float fs[27];
#pragma unroll
    for (int i = 0; i < 27; ++i) fs[i] = (float)(f[i] - W27[i]);
    ao = collide27_shifted(fs, a.beta, ac, 2.0, rho, u);
    if (active) {
#pragma unroll
      for (int i = 0; i < 27; ++i) static_cast<float*>(fout_)[(int64_t)i * n + k] = fs[i];
    }
    if (active && keep_alpha) alpha[k] = ao.alpha;
    queue_far(a, kFarShards, active && ao.status == ALPHA_DEFER, (unsigned)k, ao.alpha);
  } else {
    ao = collide27_fp64(f, a.beta, ac, 2.0, rho, u);
    if (active) {
#pragma unroll
      for (int i = 0; i < 27; ++i) static_cast<double*>(fout_)[(int64_t)i * n + k] = f[i];
    }
    if (active && keep_alpha) alpha[k] = ao.alpha;
  }
  wave_stats(stats, active, ao.iters, ao.status == ALPHA_DEFER ? -1 : ao.status, false,
             ao.logpath);
  if (active && nibb) {
    unsigned long long* sl = stats + (size_t)((blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64) &
                                              (kStatSlots - 1)) * kStatFields;
    atomicAdd(sl + ST_IBB, (unsigned long long)nibb);
  }
}

// uniform state: fluid nodes feq(rho, u), solid / null nodes the rest state w
template <bool kShifted>
__global__ void k27_init_uniform(void* __restrict__ dst, Box b, const uint8_t* __restrict__ flags,
                                 double rho, double ux, double uy, double uz) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= b.n) return;
  const bool solid = flags && flags[k] == 2;
  double u[3] = {solid ? 0.0 : ux, solid ? 0.0 : uy, solid ? 0.0 : uz};
  double feq[27];
  feq27_fp64(solid ? 1.0 : rho, u, feq);
#pragma unroll
  for (int i = 0; i < 27; ++i) {
    if constexpr (kShifted) static_cast<float*>(dst)[(int64_t)i * b.n + k] = (float)(feq[i] - W27[i]);
    else static_cast<double*>(dst)[(int64_t)i * b.n + k] = feq[i];
  }
}

inline unsigned blocks_for(int64_t n, int t) { return (unsigned)((n + t - 1) / t); }

}  // namespace

hipError_t launch_d3q27_step_periodic(const Box& b, const void* fin, void* fout, double* alpha,
                                      int keep_alpha, unsigned long long* stats,
                                      const StepArgs& a, hipStream_t s) {
  const unsigned g = blocks_for(b.n, 256);
  const uint8_t* nf = nullptr;
  if (a.shifted) {
    if (keep_alpha) k27_step_periodic<true, true, false><<<g, 256, 0, s>>>(fin, fout, alpha, stats, b, a, nf, BcTables{});
    else k27_step_periodic<true, false, false><<<g, 256, 0, s>>>(fin, fout, alpha, stats, b, a, nf, BcTables{});
  } else {
    if (keep_alpha) k27_step_periodic<false, true, false><<<g, 256, 0, s>>>(fin, fout, alpha, stats, b, a, nf, BcTables{});
    else k27_step_periodic<false, false, false><<<g, 256, 0, s>>>(fin, fout, alpha, stats, b, a, nf, BcTables{});
  }
  return hipGetLastError();
}

hipError_t launch_d3q27_step_links(const Box& b, const void* fin, void* fout, double* alpha,
                                   int keep_alpha, unsigned long long* stats, const StepArgs& a,
                                   const BcArgs& bc, hipStream_t s) {
  const unsigned g = blocks_for(b.n, 256);
  const BcTables t{bc.feq_in, bc.feq_out, bc.feq_fs};
  if (a.shifted) {
    if (keep_alpha) k27_step_periodic<true, true, true><<<g, 256, 0, s>>>(fin, fout, alpha, stats, b, a, bc.flags, t);
    else k27_step_periodic<true, false, true><<<g, 256, 0, s>>>(fin, fout, alpha, stats, b, a, bc.flags, t);
  } else {
    if (keep_alpha) k27_step_periodic<false, true, true><<<g, 256, 0, s>>>(fin, fout, alpha, stats, b, a, bc.flags, t);
    else k27_step_periodic<false, false, true><<<g, 256, 0, s>>>(fin, fout, alpha, stats, b, a, bc.flags, t);
  }
  return hipGetLastError();
}

hipError_t launch_d3q27_step_boxes(unsigned nn, const void* fin, void* fout, double* alpha,
                                   int keep_alpha, unsigned long long* stats, const StepArgs& a,
                                   const uint8_t* flags, const uint4* alias, const BoxTile* tiles,
                                   unsigned ntiles, const BoxDesc* boxes, int tile_nodes,
                                   int links, hipStream_t s) {
  if (!ntiles) return hipSuccess;
  if (tile_nodes < 64 || tile_nodes > 256 || tile_nodes % 64) return hipErrorInvalidValue;
  const dim3 g(ntiles), blk((unsigned)tile_nodes);
#define UCLBS_K27B(SH, KA, LK) \
  k27_step_boxes<SH, KA, LK><<<g, blk, 0, s>>>(fin, fout, alpha, stats, nn, a, flags, alias, tiles, boxes, tile_nodes)
#define UCLBS_K27B_LK(SH, KA) \
  do { if (links) UCLBS_K27B(SH, KA, true); else UCLBS_K27B(SH, KA, false); } while (0)
  if (a.shifted) {
    if (keep_alpha) UCLBS_K27B_LK(true, true); else UCLBS_K27B_LK(true, false);
  } else {
    if (keep_alpha) UCLBS_K27B_LK(false, true); else UCLBS_K27B_LK(false, false);
  }
#undef UCLBS_K27B_LK
#undef UCLBS_K27B
  return hipGetLastError();
}

hipError_t launch_d3q27_boundary(const Box& b, const void* fin, void* fout, double* alpha,
                                 int keep_alpha, unsigned long long* stats, const StepArgs& a,
                                 const BcArgs& bc, hipStream_t s) {
  if (bc.nlist == 0) return hipSuccess;
  const unsigned gb = blocks_for(bc.nlist, 64);
  if (a.shifted) k27_boundary<true><<<gb, 64, 0, s>>>(fin, fout, alpha, keep_alpha, stats, b, a, bc);
  else k27_boundary<false><<<gb, 64, 0, s>>>(fin, fout, alpha, keep_alpha, stats, b, a, bc);
  return hipGetLastError();
}

hipError_t launch_d3q27_init_uniform(const Box& b, void* f, int shifted, const uint8_t* flags,
                                     double rho, double ux, double uy, double uz, hipStream_t s) {
  if (shifted) k27_init_uniform<true><<<blocks_for(b.n, 256), 256, 0, s>>>(f, b, flags, rho, ux, uy, uz);
  else k27_init_uniform<false><<<blocks_for(b.n, 256), 256, 0, s>>>(f, b, flags, rho, ux, uy, uz);
  return hipGetLastError();
}

hipError_t launch_d3q27_far(const Box& b, void* fout, double* alpha, int keep_alpha,
                            unsigned long long* stats, const StepArgs& a, hipStream_t s) {
  // up to 32768 one-wave blocks (enough for a 2M-node queue at one node per
  // lane), grid-stride beyond; blocks past the queue's end leave at once.
  // Measured on the C2 start-up: 2,560 blocks 350 us, 8,192 312 us, 32,768
  // 304 us per launch.
  const unsigned g = (unsigned)std::min<int64_t>(blocks_for(b.n, kFarBlock), 32768);
  k27_far<<<g, kFarBlock, 0, s>>>(static_cast<float*>(fout), alpha, stats, b, a, keep_alpha);
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

hipError_t launch_d3q27_probe(const void* f, int64_t n, int shifted, const uint32_t* nodes,
                              unsigned m, double* out, hipStream_t s) {
  if (m == 0) return hipSuccess;
  if (shifted) k27_probe<true><<<blocks_for(m, 64), 64, 0, s>>>(f, n, nodes, m, out);
  else k27_probe<false><<<blocks_for(m, 64), 64, 0, s>>>(f, n, nodes, m, out);
  return hipGetLastError();
}

hipError_t launch_d3q27_init_tgv_region(const Box& dom, void* f, int64_t nn, int64_t off,
                                        const int ext[3], const int org[3], int shifted, double u0,
                                        double noise_amp, uint64_t seed, hipStream_t s) {
  const int64_t n = (int64_t)ext[0] * ext[1] * ext[2];
  if (shifted)
    k27_init_tgv<true><<<blocks_for(n, 256), 256, 0, s>>>(f, dom, nn, off, ext[0], ext[1], ext[2],
                                                           org[0], org[1], org[2], u0, noise_amp, seed);
  else
    k27_init_tgv<false><<<blocks_for(n, 256), 256, 0, s>>>(f, dom, nn, off, ext[0], ext[1], ext[2],
                                                            org[0], org[1], org[2], u0, noise_amp, seed);
  return hipGetLastError();
}

hipError_t launch_d3q27_init_tgv(const Box& b, void* f, int shifted, double u0,
                                 double noise_amp, uint64_t seed, hipStream_t s) {
  if (shifted)
    k27_init_tgv<true><<<blocks_for(b.n, 256), 256, 0, s>>>(f, b, b.n, 0, b.nx, b.ny, b.nz, 0, 0, 0,
                                                             u0, noise_amp, seed);
  else
    k27_init_tgv<false><<<blocks_for(b.n, 256), 256, 0, s>>>(f, b, b.n, 0, b.nx, b.ny, b.nz, 0, 0, 0,
                                                              u0, noise_amp, seed);
  return hipGetLastError();
}

}  // namespace uclbs
