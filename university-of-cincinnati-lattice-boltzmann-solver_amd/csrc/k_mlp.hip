// Learned initial guess of the D3Q27 entropic alpha: feature logging,
// hand-written training kernels (FP32 forward / backward / Adam) and the
// inference kernel on BF16 MFMA with FP32 accumulation.
//
// The network (12 -> 64 -> 64 -> 64 -> 1, GELU, 9,217 parameters) maps 12
// scale-free invariants of a node's non-equilibrium part to
// y = (alpha - 2) / m, m = max_i |D_i / f_i|; the Newton root-find of
// H(f + alpha D) = H(f) then starts from alpha0 = 2 + m (y_mean + y_std * net)
// and falls back to the plain start alpha = 2 when that start does not
// converge.  The reference has no learned component (its closest precedent,
// the fitted multiplier guesses of amr_processes.f95:1871-2039, is disabled at
// calculate_lagrangian_mults.f95:729-741).
//
// Parameter layout (FP32 master copy): W1 (12 x 64, row-major in x out), b1,
// W2 (64 x 64), b2, W3 (64 x 64), b3, W4 (64), b4.
#include "d3q27.h"
#include "kernels.h"
#include "philox.h"

#include <hip/hip_bf16.h>

namespace uclbs {

namespace {

constexpr int kIn = 12, kH = 64;
constexpr int oW1 = 0, oB1 = oW1 + kIn * kH, oW2 = oB1 + kH, oB2 = oW2 + kH * kH,
              oW3 = oB2 + kH, oB3 = oW3 + kH * kH, oW4 = oB3 + kH, oB4 = oW4 + kH;
static_assert(oB4 + 1 == kMlpParams, "parameter count");

__device__ __forceinline__ float gelu(float x) {
  const float c = 0.7978845608028654f;  // sqrt(2 / pi)
  return 0.5f * x * (1.0f + tanhf(c * (x + 0.044715f * x * x * x)));
}
__device__ __forceinline__ float gelu_grad(float x) {
  const float c = 0.7978845608028654f;
  const float t = tanhf(c * (x + 0.044715f * x * x * x));
  return 0.5f * (1.0f + t) + 0.5f * x * (1.0f - t * t) * c * (1.0f + 3.0f * 0.044715f * x * x);
}

// The 12 features of a node from its pre-collision shifted populations.
// Returns m = max|x| (0 for a node at equilibrium or with f <= 0).
__device__ __forceinline__ double features27(const float (&f)[27], float (&phi)[kIn]) {
  float D[27];
  double rho, u[3];
  prep27_shifted(f, D, rho, u);
  float m = 0.0f, xmin = 0.0f;
  bool nonpos = false;
  double S[7] = {0, 0, 0, 0, 0, 0, 0};
#pragma unroll
  for (int i = 0; i < 27; ++i) {
    const float fi = (float)W27[i] + f[i];
    nonpos |= !(fi > 0.0f);
    const float x = D[i] / fi;
    m = fmaxf(m, fabsf(x));
    xmin = fminf(xmin, x);
    float t = D[i] * x;
    S[2] += (double)t;
#pragma unroll
    for (int n = 3; n <= 6; ++n) {
      t *= x;
      S[n] += (double)t;
    }
  }
  if (nonpos || !(m > 1e-30f) || !(S[2] > 0.0)) {
#pragma unroll
    for (int j = 0; j < kIn; ++j) phi[j] = 0.0f;
    return 0.0;
  }
  const double md = m;
  phi[0] = (float)(log10(md) * 0.25 + 1.0);
  double mp = 1.0;
#pragma unroll
  for (int n = 3; n <= 6; ++n) {
    mp *= md;
    phi[n - 2] = (float)(S[n] / S[2] / mp);
  }
  phi[5] = xmin / m;
  phi[6] = (float)(S[2] / (rho * md * md));
  phi[7] = (float)(5.0 * u[0]);
  phi[8] = (float)(5.0 * u[1]);
  phi[9] = (float)(5.0 * u[2]);
  phi[10] = (float)(25.0 * (u[0] * u[0] + u[1] * u[1] + u[2] * u[2]));
  phi[11] = (float)(log10(S[2]) * 0.125 + 1.0);
  return md;
}

// pull of an interior node of a periodic box or of a level whose flag is
// kFlagInterior (no boundary link): f~_i(x) = f~_i(x - c_i) with wrap
__device__ __forceinline__ void pull27(const float* __restrict__ fin, unsigned node, const Box& b,
                                       float (&f)[27]) {
  const unsigned nx = (unsigned)b.nx, ny = (unsigned)b.ny, nz = (unsigned)b.nz, nn = (unsigned)b.n;
  const unsigned x = node % nx, yz = node / nx, y = yz % ny, z = yz / ny;
  unsigned ox[3], oy[3], oz[3];
  ox[0] = x + 1 == nx ? 0u : x + 1;
  ox[1] = x;
  ox[2] = x == 0 ? nx - 1 : x - 1;
  oy[0] = (y + 1 == ny ? 0u : y + 1) * nx;
  oy[1] = y * nx;
  oy[2] = (y == 0 ? ny - 1 : y - 1) * nx;
  oz[0] = (z + 1 == nz ? 0u : z + 1) * nx * ny;
  oz[1] = z * nx * ny;
  oz[2] = (z == 0 ? nz - 1 : z - 1) * nx * ny;
#pragma unroll
  for (int i = 0; i < 27; ++i)
    f[i] = (fin + (size_t)i * nn)[ox[CX27[i] + 1] + oy[CY27[i] + 1] + oz[CZ27[i] + 1]];
}

// Sample logging: every stride-th interior node with m above 1e-8 appends
// (features, node, m) to the sample buffer; the target is filled after the
// plain step has produced alpha (k_collect_targets).
__global__ __launch_bounds__(256) void k27_collect(const float* __restrict__ fin, Box b,
                                                   const uint8_t* __restrict__ flags, int stride,
                                                   int phase, float* __restrict__ X,
                                                   unsigned* __restrict__ node_of,
                                                   float* __restrict__ m_of,
                                                   unsigned long long* __restrict__ count,
                                                   unsigned long long cap) {
  const unsigned gid = blockIdx.x * blockDim.x + threadIdx.x;
  const unsigned nn = (unsigned)b.n;
  bool take = gid < nn && (int)(gid % (unsigned)stride) == phase;
  if (take && flags) take = flags[gid] == kFlagInterior;
  float f[27], phi[kIn];
  double m = 0.0;
  if (take) {
    pull27(fin, gid, b, f);
    m = features27(f, phi);
    take = m > 1e-8;
  }
  const unsigned long long bm = __ballot(take);
  if (!bm) return;
  unsigned long long base = 0;
  const int leader = __ffsll((long long)bm) - 1;
  if ((int)__lane_id() == leader) base = atomicAdd(count, (unsigned long long)__popcll(bm));
  base = __shfl(base, leader);
  const unsigned long long slot = base + __popcll(bm & ((1ull << __lane_id()) - 1ull));
  if (take && slot < cap) {
#pragma unroll
    for (int j = 0; j < kIn; ++j) X[slot * kIn + j] = phi[j];
    node_of[slot] = gid;
    m_of[slot] = (float)m;
  }
}

__global__ void k_collect_targets(const double* __restrict__ alpha, const unsigned* __restrict__ node_of,
                                  const float* __restrict__ m_of, float* __restrict__ y,
                                  unsigned long long first, unsigned long long last) {
  const unsigned long long k = first + blockIdx.x * (unsigned long long)blockDim.x + threadIdx.x;
  if (k >= last) return;
  y[k] = (float)((alpha[node_of[k]] - 2.0) / (double)m_of[k]);
}

// ---- training (FP32) -------------------------------------------------------
// Z[b][o] = sum_k A[b][k] W[k][o] + bias[o]; Aout = gelu(Z) (or Z)
__global__ void k_fwd_layer(const float* __restrict__ A, int K, const float* __restrict__ W,
                            const float* __restrict__ bias, int O, int B, int act,
                            float* __restrict__ Z, float* __restrict__ Aout) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= B * O) return;
  const int r = t / O, o = t % O;
  float s = bias[o];
  for (int k = 0; k < K; ++k) s = fmaf(A[r * K + k], W[k * O + o], s);
  Z[t] = s;
  Aout[t] = act ? gelu(s) : s;
}

// dZ of the output layer and the batch loss: dZ = (yhat - y) / B, loss = 0.5 mean r^2
__global__ void k_out_grad(const float* __restrict__ yhat, const float* __restrict__ y,
                           const unsigned* __restrict__ idx, int B, float* __restrict__ dZ,
                           double* __restrict__ loss) {
  __shared__ double red[256];
  double acc = 0.0;
  for (int r = threadIdx.x; r < B; r += blockDim.x) {
    const float yt = idx ? y[idx[r]] : y[r];
    const float d = yhat[r] - yt;
    dZ[r] = d / (float)B;
    acc += 0.5 * (double)d * d;
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) *loss = red[0] / B;
}

// dW[k][o] = sum_b A[b][k] dZ[b][o], db[o] = sum_b dZ[b][o]
__global__ void k_bwd_weights(const float* __restrict__ A, int K, const float* __restrict__ dZ,
                              int O, int B, float* __restrict__ gW, float* __restrict__ gb) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < K * O) {
    const int k = t / O, o = t % O;
    float s = 0.0f;
    for (int r = 0; r < B; ++r) s = fmaf(A[r * K + k], dZ[r * O + o], s);
    gW[t] = s;
  } else if (t < K * O + O) {
    const int o = t - K * O;
    float s = 0.0f;
    for (int r = 0; r < B; ++r) s += dZ[r * O + o];
    gb[o] = s;
  }
}

// dZprev[b][k] = (sum_o dZ[b][o] W[k][o]) * gelu'(Zprev[b][k])
__global__ void k_bwd_input(const float* __restrict__ dZ, int O, const float* __restrict__ W,
                            const float* __restrict__ Zprev, int K, int B,
                            float* __restrict__ dZprev) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= B * K) return;
  const int r = t / K, k = t % K;
  float s = 0.0f;
  for (int o = 0; o < O; ++o) s = fmaf(dZ[r * O + o], W[k * O + o], s);
  dZprev[t] = s * gelu_grad(Zprev[t]);
}

__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, int n, float lr, float c1, float c2) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const float gk = g[k];
  const float mk = 0.9f * m[k] + 0.1f * gk;
  const float vk = 0.999f * v[k] + 0.001f * gk * gk;
  m[k] = mk;
  v[k] = vk;
  p[k] -= lr * (mk / c1) / (sqrtf(vk / c2) + 1e-8f);
}

// gather a random batch of samples (Philox4x32-10 indices, with replacement)
__global__ void k_gather(const float* __restrict__ X, unsigned long long n, uint64_t seed,
                         uint64_t step, int B, float* __restrict__ Xb, unsigned* __restrict__ idx) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= B) return;
  uint32_t c[4] = {(uint32_t)r, (uint32_t)step, (uint32_t)(step >> 32), 0x6d6c70u};
  philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
  const unsigned long long k = (((unsigned long long)c[0] << 32) | c[1]) % n;
  idx[r] = (unsigned)k;
#pragma unroll
  for (int j = 0; j < kIn; ++j) Xb[r * kIn + j] = X[k * kIn + j];
}

// mean and standard deviation of the targets (one block, FP64)
__global__ void k_moments(const float* __restrict__ y, unsigned long long n, double* __restrict__ out) {
  __shared__ double s1[256], s2[256];
  double a = 0.0, b = 0.0;
  for (unsigned long long k = threadIdx.x; k < n; k += blockDim.x) {
    a += y[k];
    b += (double)y[k] * y[k];
  }
  s1[threadIdx.x] = a;
  s2[threadIdx.x] = b;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      s1[threadIdx.x] += s1[threadIdx.x + s];
      s2[threadIdx.x] += s2[threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double mean = s1[0] / (double)n;
    out[0] = mean;
    out[1] = sqrt(fmax(s2[0] / (double)n - mean * mean, 1e-30));
  }
}

// y <- (y - mean) / std on the device copy the trainer uses
__global__ void k_normalise(float* __restrict__ y, unsigned long long n, const double* __restrict__ ms) {
  const unsigned long long k = blockIdx.x * (unsigned long long)blockDim.x + threadIdx.x;
  if (k < n) y[k] = (float)(((double)y[k] - ms[0]) / ms[1]);
}

// ---- inference on BF16 MFMA -----------------------------------------------
// One wavefront evaluates 64 nodes.  v_mfma_f32_16x16x16_bf16 layouts (wave64):
// A (16x16): lane l holds row l%16, k = 4(l/16)..4(l/16)+3; B (16x16): lane l
// holds column l%16, the same k; D (16x16 FP32): lane l holds column l%16,
// rows 4(l/16)..4(l/16)+3.
typedef __attribute__((__vector_size__(4 * sizeof(short)))) short bf16x4;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4;

__device__ __forceinline__ short bf16_bits(float x) {
  return (short)__bfloat16_as_ushort(__float2bfloat16(x));
}

// one dense layer for the wave's 64 rows: out[64][64] = act(in[64][K] W[K][64] + b),
// in/out in LDS as bf16 bits with a row pitch of 72 (bank spread)
constexpr int kPitch = 72;
template <int K>
__device__ __forceinline__ void mfma_layer(const short* __restrict__ in, const short* __restrict__ Wt,
                                           const float* __restrict__ bias, short* __restrict__ out,
                                           bool act, float* __restrict__ out_f) {
  const int l = __lane_id();
  const int r16 = l % 16, kq = 4 * (l / 16);
#pragma unroll
  for (int rt = 0; rt < 4; ++rt) {
#pragma unroll
    for (int ct = 0; ct < 4; ++ct) {
      f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int kk = 0; kk < K / 16; ++kk) {
        bf16x4 a, bv;
        const short* ap = in + (rt * 16 + r16) * kPitch + kk * 16 + kq;
        // Wt is W transposed: Wt[o][k], so a lane's 4 k-values are contiguous
        const short* bp = Wt + (ct * 16 + r16) * K + kk * 16 + kq;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          a[j] = ap[j];
          bv[j] = bp[j];
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, bv, acc, 0, 0, 0);
      }
      const int col = ct * 16 + r16;
      const float bb = bias[col];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int row = rt * 16 + kq + j;
        const float z = acc[j] + bb;
        const float v = act ? gelu(z) : z;
        if (out) out[row * kPitch + col] = bf16_bits(v);
        if (out_f) out_f[row * kPitch + col] = v;
      }
    }
  }
}

// Network output for the wave's 64 rows of phi (one per lane); returns the
// lane's y (normalised).  Wt1: 64 x 16 (K padded), Wt2, Wt3: 64 x 64, all
// bf16 bits, transposed; bias / W4 / b4 FP32 from the master copy.
__device__ __forceinline__ float mlp_wave(const float (&phi)[kIn], const short* __restrict__ Wt1,
                                          const short* __restrict__ Wt2, const short* __restrict__ Wt3,
                                          const float* __restrict__ P, short* __restrict__ bufA,
                                          short* __restrict__ bufB, float* __restrict__ bufF) {
  const int l = __lane_id();
#pragma unroll
  for (int j = 0; j < 16; ++j) bufA[l * kPitch + j] = bf16_bits(j < kIn ? phi[j] : 0.0f);
  __builtin_amdgcn_wave_barrier();
  mfma_layer<16>(bufA, Wt1, P + oB1, bufB, true, nullptr);
  __builtin_amdgcn_wave_barrier();
  mfma_layer<64>(bufB, Wt2, P + oB2, bufA, true, nullptr);
  __builtin_amdgcn_wave_barrier();
  mfma_layer<64>(bufA, Wt3, P + oB3, nullptr, true, bufF);
  __builtin_amdgcn_wave_barrier();
  float y = P[oB4];
#pragma unroll 8
  for (int k = 0; k < kH; ++k) y = fmaf(bufF[l * kPitch + k], P[oW4 + k], y);
  return y;
}

// forward of a host-given batch through the MFMA path (checks of the layout)
__global__ __launch_bounds__(64) void k_mlp_infer(const float* __restrict__ X, int B,
                                                  const short* __restrict__ Wbf,
                                                  const float* __restrict__ P,
                                                  float* __restrict__ yhat) {
  __shared__ short bufA[64 * kPitch], bufB[64 * kPitch];
  __shared__ float bufF[64 * kPitch];
  const int r = blockIdx.x * 64 + threadIdx.x;
  float phi[kIn];
#pragma unroll
  for (int j = 0; j < kIn; ++j) phi[j] = r < B ? X[r * kIn + j] : 0.0f;
  const float y = mlp_wave(phi, Wbf, Wbf + 64 * 16, Wbf + 64 * 16 + 64 * 64, P, bufA, bufB, bufF);
  if (r < B) yhat[r] = y;
}

// Initial guesses for the step: interior nodes get alpha0 = 2 + m (mean +
// std * net(phi)); every other node 2.  Written to alpha[].
__global__ __launch_bounds__(64) void k27_alpha_guess(const float* __restrict__ fin, Box b,
                                                      const uint8_t* __restrict__ flags,
                                                      const short* __restrict__ Wbf,
                                                      const float* __restrict__ P, float ymean,
                                                      float ystd, double* __restrict__ alpha) {
  __shared__ short bufA[64 * kPitch], bufB[64 * kPitch];
  __shared__ float bufF[64 * kPitch];
  const unsigned gid = blockIdx.x * 64u + threadIdx.x;
  const unsigned nn = (unsigned)b.n;
  bool ok = gid < nn;
  if (ok && flags) ok = flags[gid] == kFlagInterior;
  float f[27], phi[kIn];
  double m = 0.0;
  if (ok) {
    pull27(fin, gid, b, f);
    m = features27(f, phi);
  } else {
#pragma unroll
    for (int j = 0; j < kIn; ++j) phi[j] = 0.0f;
  }
  // the whole wave evaluates the network (MFMA needs all lanes)
  const float y = mlp_wave(phi, Wbf, Wbf + 64 * 16, Wbf + 64 * 16 + 64 * 64, P, bufA, bufB, bufF);
  if (gid < nn) alpha[gid] = m > 1e-12 ? 2.0 + m * ((double)ymean + (double)ystd * (double)y) : 2.0;
}

// bf16 transposed copies of W1 (padded to K = 16), W2, W3 for the MFMA path
__global__ void k_pack_bf16(const float* __restrict__ P, short* __restrict__ Wbf) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < 64 * 16) {
    const int o = t / 16, k = t % 16;
    Wbf[t] = bf16_bits(k < kIn ? P[oW1 + k * kH + o] : 0.0f);
  } else if (t < 64 * 16 + 2 * 64 * 64) {
    const int u = t - 64 * 16;
    const int layer = u / 4096, v = u % 4096;
    const int o = v / 64, k = v % 64;
    Wbf[t] = bf16_bits(P[(layer == 0 ? oW2 : oW3) + k * kH + o]);
  }
}

}  // namespace

hipError_t launch_mlp_collect(const Box& b, const float* fin, const uint8_t* flags, int stride,
                              int phase, MlpSamples& s, hipStream_t st) {
  const unsigned blocks = (unsigned)((b.n + 255) / 256);
  hipLaunchKernelGGL(k27_collect, dim3(blocks), dim3(256), 0, st, fin, b, flags, stride, phase,
                     s.X, s.node_of, s.m_of, s.count, (unsigned long long)s.cap);
  return hipGetLastError();
}

hipError_t launch_mlp_targets(const double* alpha, MlpSamples& s, unsigned long long first,
                              unsigned long long last, hipStream_t st) {
  if (last <= first) return hipSuccess;
  const unsigned blocks = (unsigned)((last - first + 255) / 256);
  hipLaunchKernelGGL(k_collect_targets, dim3(blocks), dim3(256), 0, st, alpha, s.node_of, s.m_of,
                     s.y, first, last);
  return hipGetLastError();
}

// forward + backward of one batch (X: B x 12, y: targets; idx maps rows into
// y or is null); gradients into g (kMlpParams), loss into *loss
hipError_t launch_mlp_loss_grad(const float* P, const float* X, const float* y,
                                const unsigned* idx, int B, MlpWork& w, float* g, double* loss,
                                hipStream_t st) {
  const int T = 256;
  auto nb = [&](int n) { return dim3((unsigned)((n + T - 1) / T)); };
  hipLaunchKernelGGL(k_fwd_layer, nb(B * kH), dim3(T), 0, st, X, kIn, P + oW1, P + oB1, kH, B, 1, w.Z1, w.A1);
  hipLaunchKernelGGL(k_fwd_layer, nb(B * kH), dim3(T), 0, st, w.A1, kH, P + oW2, P + oB2, kH, B, 1, w.Z2, w.A2);
  hipLaunchKernelGGL(k_fwd_layer, nb(B * kH), dim3(T), 0, st, w.A2, kH, P + oW3, P + oB3, kH, B, 1, w.Z3, w.A3);
  hipLaunchKernelGGL(k_fwd_layer, nb(B), dim3(T), 0, st, w.A3, kH, P + oW4, P + oB4, 1, B, 0, w.Z4, w.A4);
  hipLaunchKernelGGL(k_out_grad, dim3(1), dim3(256), 0, st, w.A4, y, idx, B, w.D4, loss);
  hipLaunchKernelGGL(k_bwd_weights, nb(kH + 1), dim3(T), 0, st, w.A3, kH, w.D4, 1, B, g + oW4, g + oB4);
  hipLaunchKernelGGL(k_bwd_input, nb(B * kH), dim3(T), 0, st, w.D4, 1, P + oW4, w.Z3, kH, B, w.D3);
  hipLaunchKernelGGL(k_bwd_weights, nb(kH * kH + kH), dim3(T), 0, st, w.A2, kH, w.D3, kH, B, g + oW3, g + oB3);
  hipLaunchKernelGGL(k_bwd_input, nb(B * kH), dim3(T), 0, st, w.D3, kH, P + oW3, w.Z2, kH, B, w.D2);
  hipLaunchKernelGGL(k_bwd_weights, nb(kH * kH + kH), dim3(T), 0, st, w.A1, kH, w.D2, kH, B, g + oW2, g + oB2);
  hipLaunchKernelGGL(k_bwd_input, nb(B * kH), dim3(T), 0, st, w.D2, kH, P + oW2, w.Z1, kH, B, w.D1);
  hipLaunchKernelGGL(k_bwd_weights, nb(kIn * kH + kH), dim3(T), 0, st, X, kIn, w.D1, kH, B, g + oW1, g + oB1);
  return hipGetLastError();
}

hipError_t launch_mlp_forward_fp32(const float* P, const float* X, int B, MlpWork& w, float* yhat,
                                   hipStream_t st) {
  const int T = 256;
  auto nb = [&](int n) { return dim3((unsigned)((n + T - 1) / T)); };
  hipLaunchKernelGGL(k_fwd_layer, nb(B * kH), dim3(T), 0, st, X, kIn, P + oW1, P + oB1, kH, B, 1, w.Z1, w.A1);
  hipLaunchKernelGGL(k_fwd_layer, nb(B * kH), dim3(T), 0, st, w.A1, kH, P + oW2, P + oB2, kH, B, 1, w.Z2, w.A2);
  hipLaunchKernelGGL(k_fwd_layer, nb(B * kH), dim3(T), 0, st, w.A2, kH, P + oW3, P + oB3, kH, B, 1, w.Z3, w.A3);
  hipLaunchKernelGGL(k_fwd_layer, nb(B), dim3(T), 0, st, w.A3, kH, P + oW4, P + oB4, 1, B, 0, w.Z4, yhat);
  return hipGetLastError();
}

hipError_t launch_mlp_adam(float* P, const float* g, float* m, float* v, int t, float lr,
                           hipStream_t st) {
  const float c1 = 1.0f - powf(0.9f, (float)t), c2 = 1.0f - powf(0.999f, (float)t);
  hipLaunchKernelGGL(k_adam, dim3((kMlpParams + 255) / 256), dim3(256), 0, st, P, g, m, v,
                     kMlpParams, lr, c1, c2);
  return hipGetLastError();
}

hipError_t launch_mlp_gather(const float* X, unsigned long long n, uint64_t seed, uint64_t step,
                             int B, float* Xb, unsigned* idx, hipStream_t st) {
  hipLaunchKernelGGL(k_gather, dim3((B + 255) / 256), dim3(256), 0, st, X, n, seed, step, B, Xb, idx);
  return hipGetLastError();
}

hipError_t launch_mlp_normalise(float* y, unsigned long long n, double* ms, hipStream_t st) {
  hipLaunchKernelGGL(k_moments, dim3(1), dim3(256), 0, st, y, n, ms);
  hipLaunchKernelGGL(k_normalise, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, y, n, ms);
  return hipGetLastError();
}

hipError_t launch_mlp_pack(const float* P, short* Wbf, hipStream_t st) {
  hipLaunchKernelGGL(k_pack_bf16, dim3((64 * 16 + 2 * 4096 + 255) / 256), dim3(256), 0, st, P, Wbf);
  return hipGetLastError();
}

hipError_t launch_mlp_infer(const float* X, int B, const short* Wbf, const float* P, float* yhat,
                            hipStream_t st) {
  hipLaunchKernelGGL(k_mlp_infer, dim3((B + 63) / 64), dim3(64), 0, st, X, B, Wbf, P, yhat);
  return hipGetLastError();
}

hipError_t launch_alpha_guess27(const Box& b, const float* fin, const uint8_t* flags,
                                const short* Wbf, const float* P, float ymean, float ystd,
                                double* alpha, hipStream_t st) {
  hipLaunchKernelGGL(k27_alpha_guess, dim3((unsigned)((b.n + 63) / 64)), dim3(64), 0, st, fin, b,
                     flags, Wbf, P, ymean, ystd, alpha);
  return hipGetLastError();
}

}  // namespace uclbs
