// D3Q27 entropic collide for one node, two storage forms.
//
// The reference has no D3Q27 collision (collision.f95:211-222 handles only
// dir 19 and 39; directions_and_weights.f95:226-237 marks dir=27 as "XXXX"), so
// the node update follows the published entropic LBM the oracle states:
//   rho = sum f, rho u = sum c f          (calculate_flow_variables.f95:201-207)
//   feq = product-form entropic equilibrium (Ansumali, Karlin, Ottinger 2003)
//   fout = f + alpha beta (feq - f)        (collision.f95:289 with q_bar = 0)
//   alpha from alpha_solve.h; beta = 1/(6 nu + 1).
//
// Storage forms:
//   Fp32Shifted: populations stored as f~ = f - w_i in FP32.  Moment sums run
//     in FP64; the equilibrium is built directly in shifted form,
//     feq~_i = w_i [(1+drho)(1+ex)(1+ey)(1+ez) - 1] with the 1-D factors
//     written as 1 + e so no O(1) quantity is ever rounded in FP32; x_i and
//     the products f x^n are FP32 (they carry the FP32 inputs' precision),
//     the entropy power sums, the mass residual and the alpha Newton are FP64.
//   Fp64: full populations in FP64, every step in FP64 (oracle mode).
#pragma once
#include <hip/hip_runtime.h>
#include "alpha_solve.h"
#include "lattice.h"

namespace uclbs {

struct AlphaCfg {
  double tol;
  int maxit;
  int mode;  // 0 = entropic Newton from 2, 1 = fixed alpha = 2 (LBGK), 2 = given initial guess
  SeriesThr thr;
};

__device__ __forceinline__ float frcp(float x) { return __builtin_amdgcn_rcpf(x); }

// 1-D factors Psi(c) = 1 + e(c) of the product-form equilibrium, c = -1, 0, +1.
// With s = sqrt(1 + 3u^2) and d = s - 1 = 3u^2 / (1 + s):
//   Psi(0) = 2 - s               -> e0 = -d
//   Psi(+1) = (2-s)(2u+s)/(1-u)  -> ep = (3u - 2ud - d^2) / (1 - u)
//   Psi(-1) = (2-s)(1-u)/(2u+s)  -> em = (-3u - 2d + ud) / (1 + 2u + d)
template <typename R>
__device__ __forceinline__ void axis_eps(R u, R& em, R& e0, R& ep) {
  const R s = sqrt(R(1) + R(3) * u * u);
  const R d = R(3) * u * u / (R(1) + s);
  e0 = -d;
  ep = (R(3) * u - R(2) * u * d - d * d) / (R(1) - u);
  em = (R(-3) * u - R(2) * d + u * d) / (R(1) + R(2) * u + d);
}
// FP32 form with hardware reciprocal / square root (about 1 ulp each; the
// factors only need relative accuracy ~1e-7 since they are O(u) deviations).
__device__ __forceinline__ void axis_eps_fast(float u, float& em, float& e0, float& ep) {
  const float uu = 3.0f * u * u;
  const float s = __builtin_amdgcn_sqrtf(1.0f + uu);
  const float d = uu * frcp(1.0f + s);
  e0 = -d;
  ep = (3.0f * u - 2.0f * u * d - d * d) * frcp(1.0f - u);
  em = (-3.0f * u - 2.0f * d + u * d) * frcp(1.0f + 2.0f * u + d);
}

// (1+a)(1+b) - 1 without forming 1+a
template <typename R>
__device__ __forceinline__ R mul1(R a, R b) { return a + b + a * b; }

// moments (FP64 sums), shifted equilibrium and D = feq~ - f~ with the mass
// projection; shared by the streaming and far-from-equilibrium collides
__device__ __forceinline__ void prep27_shifted(const float (&f)[27], float (&D)[27],
                                               double& rho_out, double (&u_out)[3]) {
  double r = 0.0, jx = 0.0, jy = 0.0, jz = 0.0;
#pragma unroll
  for (int i = 0; i < 27; ++i) {
    const double fi = f[i];
    r += fi;
    if (CX27[i] > 0) jx += fi; else if (CX27[i] < 0) jx -= fi;
    if (CY27[i] > 0) jy += fi; else if (CY27[i] < 0) jy -= fi;
    if (CZ27[i] > 0) jz += fi; else if (CZ27[i] < 0) jz -= fi;
  }
  const double rho = 1.0 + r;
  const double inv = 1.0 / rho;
  const double ux = jx * inv, uy = jy * inv, uz = jz * inv;
  rho_out = rho;
  u_out[0] = ux; u_out[1] = uy; u_out[2] = uz;

  float ex[3], ey[3], ez[3];
  axis_eps_fast((float)ux, ex[0], ex[1], ex[2]);
  axis_eps_fast((float)uy, ey[0], ey[1], ey[2]);
  axis_eps_fast((float)uz, ez[0], ez[1], ez[2]);
  const float dr = (float)r;
  float qx[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) qx[c] = mul1(dr, ex[c]);
  float pyz[9];
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = 0; b < 3; ++b) pyz[a * 3 + b] = mul1(ey[a], ez[b]);

  // D = feq~ - f~.  Mass projection: in exact arithmetic sum D = 0; the FP32
  // rounding of the 1-D factors leaves a small, partly systematic residual
  // (it drifted the total mass by ~1e-10 per node and step), so the residual
  // is summed in FP64 and removed through the rest population.
  double sd = 0.0;
#pragma unroll
  for (int i = 0; i < 27; ++i) {
    const float e = mul1(qx[CX27[i] + 1], pyz[(CY27[i] + 1) * 3 + CZ27[i] + 1]);
    D[i] = (float)W27[i] * e - f[i];
    sd += (double)D[i];
  }
  D[0] -= (float)sd;
}

// ---------------------------------------------------------------- FP32 shifted
// f[i] holds f~_i on entry and fout~_i on return.  A node whose max|x| lies
// beyond the in-register series tiers returns ALPHA_DEFER with f[] untouched
// and, in `alpha`, the root of the order-kMidOrder truncated series as a
// starting point; the caller stores the pre-collision state and queues the
// node for the far-from-equilibrium kernel, so one far lane does not stall
// its wavefront.
__device__ __forceinline__ AlphaOut collide27_shifted(float (&f)[27], double beta,
                                                      const AlphaCfg& ac, double alpha0,
                                                      double& rho_out, double (&u_out)[3]) {
  float D[27];
  prep27_shifted(f, D, rho_out, u_out);

  AlphaOut ao{2.0, 0, ALPHA_TRIVIAL, false};
  if (ac.mode != 1) {
    // One fused pass: x_i = D_i / f_i (FP32, hardware reciprocal), max|x|,
    // min x, and the power sums S_n = sum f x^n = sum D x^(n-1), n = 2..6:
    // FP32 products summed in FP32 over three groups of nine directions, the
    // group sums in FP64 (the summation adds errors of the size the FP32
    // terms already carry, and no O(1) quantity is summed in FP32).
    float m = 0.0f, xmin = 0.0f;
    bool nonpos = false;
    double S[kFastOrder + 1];
#pragma unroll
    for (int n = 2; n <= kFastOrder; ++n) S[n] = 0.0;
#pragma unroll
    for (int g = 0; g < 3; ++g) {
      float P[kFastOrder + 1];
#pragma unroll
      for (int n = 2; n <= kFastOrder; ++n) P[n] = 0.0f;
#pragma unroll
      for (int i = 9 * g; i < 9 * g + 9; ++i) {
        const float fi = (float)W27[i] + f[i];
        nonpos |= !(fi > 0.0f);
        const float x = D[i] * frcp(fi);
        m = fmaxf(m, fabsf(x));
        xmin = fminf(xmin, x);
        float t = D[i] * x;
        P[2] += t;
#pragma unroll
        for (int n = 3; n <= kFastOrder; ++n) { t *= x; P[n] += t; }
      }
#pragma unroll
      for (int n = 2; n <= kFastOrder; ++n) S[n] += (double)P[n];
    }
    const double amax = xmin < 0.0f ? -1.0 / (double)xmin : 1e300;
    if (nonpos) {
      ao = AlphaOut{2.0, 0, ALPHA_NONPOS, false};
    } else if ((double)m < kTrivialX) {
      ao = AlphaOut{2.0, 0, ALPHA_TRIVIAL, false};
    } else if (amax < 2.0) {
      ao = AlphaOut{amax, 0, ALPHA_BOUND, false};
    } else if (!__any(m > ac.thr.m_fast)) {
      // the whole wavefront is within the fused order: Newton on the scalar series
      auto ev = [&](double a, double& g, double& gp) __attribute__((always_inline)) {
        series_g<kFastOrder>(S, a, g, gp);
      };
      if (ac.mode == 2) {
        // learned start; the plain start takes over when it is outside the
        // bracket or does not converge
        const bool inside = alpha0 > 1.0 && alpha0 < fmin(amax, 1e6);
        ao = newton_alpha(ev, S[2], amax, inside ? alpha0 : 2.0, ac.tol, ac.maxit);
        if (!inside || ao.status == ALPHA_NOCONV) {
          const int spent = inside ? ao.iters : 0;
          ao = newton_alpha(ev, S[2], amax, 2.0, ac.tol, ac.maxit);
          ao.iters += spent;
          ao.fallback = true;
        }
      } else {
        ao = newton_alpha(ev, S[2], amax, 2.0, ac.tol, ac.maxit);
      }
    } else {
      // some lane needs more terms: the wavefront extends its sums to order
      // kMidOrder from the populations still in registers.  Orders above 6
      // are accumulated in FP32: at max|x| <= m_mid their terms of g are below
      // 1e-6 S2, so FP32 sums leave errors far below tol * S2.
      float Sh[kMidOrder > kFastOrder ? kMidOrder - kFastOrder : 1];
#pragma unroll
      for (int q = 0; q < kMidOrder - kFastOrder; ++q) Sh[q] = 0.0f;
#pragma unroll
      for (int i = 0; i < 27; ++i) {
        const float x = D[i] * frcp((float)W27[i] + f[i]);
        const float x2 = x * x;
        float t = D[i] * x * (x2 * x2);  // D x^5, the order-6 term
#pragma unroll
        for (int q = 0; q < kMidOrder - kFastOrder; ++q) { t *= x; Sh[q] += t; }
      }
      double Shd[kMidOrder > kFastOrder ? kMidOrder - kFastOrder : 1];
#pragma unroll
      for (int q = 0; q < kMidOrder - kFastOrder; ++q) Shd[q] = (double)Sh[q];
      // one Newton for both kinds of lane: to the tolerance within the tier,
      // to a seed (1e-6, at most 8 updates) for a lane beyond it
      const bool far = m > ac.thr.m_mid;
      const double a0 = ac.mode == 2 ? alpha0 : 2.0;
      ao = newton_alpha(
          [&](double a, double& g, double& gp) __attribute__((always_inline)) {
            series_g2<kFastOrder, kMidOrder>(S, Shd, a, g, gp);
          },
          S[2], amax, a0, far ? 1e-6 : ac.tol, far ? 8 : ac.maxit);
      if (far) return AlphaOut{ao.alpha, ao.iters, ALPHA_DEFER, false};
    }
  }
  const float ab = (float)(ao.alpha * beta);
#pragma unroll
  for (int i = 0; i < 27; ++i) f[i] = fmaf(ab, D[i], f[i]);
  return ao;
}

// Far-from-equilibrium node (queued by the streaming kernel): the re-centred
// logarithmic solve of alpha_solve.h from the series seed, then relaxation.
// Same moments, equilibrium and D as the streaming kernel formed for it.
// f~ and D are parked in the block's LDS columns (stride kFarBlock = blockDim)
// so the sweeps over the 27 directions run as loops unrolled by three (three
// independent logarithms in flight) with few live registers; the relaxed
// populations are written to fout~ at node.
constexpr int kRecentreOrder = 6;
constexpr int kFarBlock = 64;
__device__ __forceinline__ AlphaOut collide27_far(float (&f)[27], double beta, const AlphaCfg& ac,
                                                  double seed, float* __restrict__ sf,
                                                  float* __restrict__ sD, bool active,
                                                  float* __restrict__ fout, unsigned nn,
                                                  unsigned node, const LogTable& lt) {
  double rho_unused, u_unused[3];
  AlphaOut ao;
  float xmin = 0.0f;
  double S2 = 0.0;
  bool nonpos = false;
  {
    float D[27];
    prep27_shifted(f, D, rho_unused, u_unused);
#pragma unroll
    for (int i = 0; i < 27; ++i) {
      const float fi = (float)W27[i] + f[i];
      nonpos |= !(fi > 0.0f);
      const float x = D[i] * frcp(fi);
      xmin = fminf(xmin, x);
      S2 += (double)(D[i] * x);
      sf[i * kFarBlock] = f[i];
      sD[i * kFarBlock] = D[i];
    }
  }
  const double amax = xmin < 0.0f ? -1.0 / (double)xmin : 1e300;
  if (nonpos) {
    ao = AlphaOut{2.0, 0, ALPHA_NONPOS, false};
  } else if (amax < 2.0) {
    ao = AlphaOut{amax, 0, ALPHA_BOUND, false};
  } else {
    auto dx = [&](int i, double& d, double& x) __attribute__((always_inline)) {
      d = (double)sD[i * kFarBlock];
      x = d * drcp(W27[i] + (double)sf[i * kFarBlock]);
    };
    auto fx = [&](int i) __attribute__((always_inline)) { return W27[i] + (double)sf[i * kFarBlock]; };
    auto lg = [&](double y) __attribute__((always_inline)) { return log1p_tab(y, lt); };
    double K = 0.0;
#pragma unroll 3
    for (int i = 0; i < 27; ++i) {
      double d, x;
      dx(i, d, x);
      K += d * lg(x);
    }
    ao = recentred_alpha<27, kRecentreOrder, 3>(dx, fx, lg, K, S2, amax, seed, ac.tol, ac.maxit);
  }
  const float ab = (float)(ao.alpha * beta);
  if (active) {
#pragma unroll
    for (int i = 0; i < 27; ++i) (fout + (size_t)i * nn)[node] = fmaf(ab, sD[i * kFarBlock], sf[i * kFarBlock]);
  }
  return ao;
}

// ---------------------------------------------------------------- FP64 full
__device__ __forceinline__ AlphaOut collide27_fp64(double (&f)[27], double beta,
                                                   const AlphaCfg& ac, double alpha0,
                                                   double& rho_out, double (&u_out)[3]) {
  double r = 0.0, jx = 0.0, jy = 0.0, jz = 0.0;
#pragma unroll
  for (int i = 0; i < 27; ++i) {
    r += f[i];
    if (CX27[i] > 0) jx += f[i]; else if (CX27[i] < 0) jx -= f[i];
    if (CY27[i] > 0) jy += f[i]; else if (CY27[i] < 0) jy -= f[i];
    if (CZ27[i] > 0) jz += f[i]; else if (CZ27[i] < 0) jz -= f[i];
  }
  const double ux = jx / r, uy = jy / r, uz = jz / r;
  rho_out = r;
  u_out[0] = ux; u_out[1] = uy; u_out[2] = uz;
  // 1-D factors in the oracle's direct form
  double px[3], py[3], pz[3];
  auto psi = [](double u, double* p) {
    const double s = sqrt(1.0 + 3.0 * u * u);
    const double b = 2.0 - s;
    p[1] = b;
    p[2] = b * (2.0 * u + s) / (1.0 - u);
    p[0] = b * (1.0 - u) / (2.0 * u + s);
  };
  psi(ux, px); psi(uy, py); psi(uz, pz);
  double D[27];
  double m = 0.0, xmin = 0.0, S2 = 0.0;
  bool nonpos = false;
#pragma unroll
  for (int i = 0; i < 27; ++i) {
    const double feq = r * W27[i] * px[CX27[i] + 1] * py[CY27[i] + 1] * pz[CZ27[i] + 1];
    D[i] = feq - f[i];
  }
  AlphaOut ao{2.0, 0, ALPHA_TRIVIAL, false};
  if (ac.mode != 1) {
#pragma unroll
    for (int i = 0; i < 27; ++i) {
      nonpos |= !(f[i] > 0.0);
      const double x = D[i] / f[i];
      m = fmax(m, fabs(x));
      xmin = fmin(xmin, x);
      S2 += D[i] * x;
    }
    if (nonpos) {
      ao = AlphaOut{2.0, 0, ALPHA_NONPOS, false};
    } else {
      auto dx = [&](int i, double& d, double& x) __attribute__((always_inline)) { d = D[i]; x = D[i] / f[i]; };
      auto fx = [&](int i) __attribute__((always_inline)) { return f[i]; };
      ao = solve_alpha<27>(dx, fx, m, xmin, S2, ac.mode == 2 ? alpha0 : 2.0, ac.tol, ac.maxit,
                           ac.thr);
    }
  }
  const double ab = ao.alpha * beta;
#pragma unroll
  for (int i = 0; i < 27; ++i) f[i] = f[i] + ab * D[i];
  return ao;
}

}  // namespace uclbs
