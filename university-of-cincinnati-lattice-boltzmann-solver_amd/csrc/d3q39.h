// D3Q39 two-population (f, g) compressible entropic node update, FP64 arithmetic.
//
// Per node and step (looping/timestepping/loop.f95:105-138 for one node):
//   moments     rho, u, T, P_ab, q_a        calculate_flow_variables.f95:201-228
//   multipliers Newton on the 13 moment constraints from the stored
//               multipliers, stop when max|delta| <= 1e-7
//                                            calculate_lagrangian_mults.f95:337-1098
//   alpha       closed-form entropic estimate from the last Newton
//               iteration's equilibrium      calculate_alpha.f95:118-289
//   collision   f' = f + alpha b1 (D - qbar) + b2 qbar,
//               g' = g + alpha b1 (Dg - gbar) + b2 gbar
//                                            collision.f95:206-293
// The equilibrium is f_eq,i = rho exp(lambda . m_i) with the monomials
// m_i = (1, c, c_x^2, c_y^2, c_z^2, c_x c_y, c_x c_z, c_y c_z, c |c|^2)
// (amr_processes.f95:1592-1747).  The Jacobian of the moments is the Gram
// matrix J = sum_i f_eq,i m_i m_i^T; it is symmetric positive definite and is
// factorised here as L D L^T in packed form, in place (the reference's
// lu_decomp.f95 performs the same elimination unsymmetrically, with two index
// slips at :105 and :441 that this library does not reproduce).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

#include "lattice.h"

namespace uclbs {

// Constants of the gas and reference state, derived on the host as the
// reference's start-up derives them (basic_fluids_calcs.f95:17-86,
// viscosity_and_timing.f95:29-73, timing_calcs.f95:24-26).
struct Gas39 {
  double gama, R, cv, cp, lead_g;     // lead_g = 2 cv - 3 (collision.f95:73)
  double t_ref, p_ref, rho_ref, mu_ref, kappa_ref, molec_weight;
  double dt, eps;                     // dt(lvl), elbm_epsilon(lvl)
  double init_beta1, init_beta2, beta1_scale, beta2_scale;
  int lvl, startup, turbulent, maxit;
};

constexpr int kM39 = 13;
constexpr int kJ39 = kM39 * (kM39 + 1) / 2;  // packed lower triangle

template <bool kShifted>
__host__ __device__ constexpr int cx39(int i) { return CX39[i] + (kShifted ? 1 : 0); }

template <bool kShifted>
__host__ __device__ constexpr int mono39(int i, int k) {
  // k-th monomial of direction i
  const int x = cx39<kShifted>(i), y = CY39[i], z = CZ39[i];
  const int c2 = x * x + y * y + z * z;
  return k == 0 ? 1 : k == 1 ? x : k == 2 ? y : k == 3 ? z : k == 4 ? x * x : k == 5 ? y * y
       : k == 6 ? z * z : k == 7 ? x * y : k == 8 ? x * z : k == 9 ? y * z : k == 10 ? x * c2
       : k == 11 ? y * c2 : z * c2;
}

__host__ __device__ constexpr int jidx(int k, int l) { return k * (k + 1) / 2 + l; }  // l <= k

template <bool kShifted>
__host__ __device__ inline double expo39(const double* __restrict__ lm, double rho, int i) {
  double e = lm[0];
#pragma unroll
  for (int k = 1; k < kM39; ++k) {
    const int m = mono39<kShifted>(i, k);
    if (m != 0) e += (double)m * lm[k];
  }
  return rho * exp(e);
}

// Jacobian entries that hold the same moment: J_kl = sum_i e_i m_k(c_i) m_l(c_i)
// depends only on the weights m_k m_l over the 39 directions, and only 59 of
// the 91 (k, l) pairs have distinct weight vectors on the shifted lattice (57
// on the plain one: x*x = 1*x^2, x^2*y^2 = xy*xy, ...).  rep[p] is the first pair of class p, cls[jidx(k, l)] its class.
struct JDup39 {
  int n[2];
  signed char rk[2][91], rl[2][91], cls[2][91];
  constexpr JDup39() : n(), rk(), rl(), cls() {
    for (int sh = 0; sh < 2; ++sh) {
      int cnt = 0;
      for (int k = 0; k < 13; ++k)
        for (int l = 0; l <= k; ++l) {
          int found = -1;
          for (int p = 0; p < cnt && found < 0; ++p) {
            bool same = true;
            for (int i = 0; i < 39 && same; ++i) {
              const int a = sh ? mono39<true>(i, k) * mono39<true>(i, l)
                               : mono39<false>(i, k) * mono39<false>(i, l);
              const int b = sh ? mono39<true>(i, rk[sh][p]) * mono39<true>(i, rl[sh][p])
                               : mono39<false>(i, rk[sh][p]) * mono39<false>(i, rl[sh][p]);
              same = a == b;
            }
            if (same) found = p;
          }
          if (found < 0) {
            rk[sh][cnt] = (signed char)k;
            rl[sh][cnt] = (signed char)l;
            found = cnt++;
          }
          cls[sh][k * (k + 1) / 2 + l] = (signed char)found;
        }
      n[sh] = cnt;
    }
  }
};
constexpr JDup39 kJDup39{};

struct NoSink39 {
  __host__ __device__ void operator()(int, double) const {}
};

// Directions grouped by which velocity components are zero (bit 0: c_x, bit 1:
// c_y, bit 2: c_z).  Half of the 39 x 59 Jacobian weights m_k m_l vanish
// because a component does (1,171 of 2,301 on the shifted lattice, 931 of 2,223
// on the plain one); within a group the vanishing ones are known at compile
// time, so the accumulation of a group skips them.  Each group is unrolled
// with its weights as literals: with the zeros gone the unrolled Newton body
// fits the instruction cache, and the weights no longer wait on scalar loads
// from a constant table, which is what bounded the kernel at one wave per
// SIMD (C3 step 8.9 -> 7.2 ms; skipping the zeros alone, with the weights
// still loaded, gave nothing).
template <bool kShifted>
__host__ __device__ constexpr int zpat39(int i) {
  return (cx39<kShifted>(i) == 0 ? 1 : 0) | (CY39[i] == 0 ? 2 : 0) | (CZ39[i] == 0 ? 4 : 0);
}

struct Pat39Table {
  int count[2][8];
  int dir[2][8][39];
  constexpr Pat39Table() : count(), dir() {
    for (int sh = 0; sh < 2; ++sh)
      for (int i = 0; i < 39; ++i) {
        const int p = sh ? zpat39<true>(i) : zpat39<false>(i);
        dir[sh][p][count[sh][p]++] = i;
      }
  }
};
constexpr Pat39Table kPatC{};

// monomial k vanishes on every direction of zero pattern pat
__host__ __device__ constexpr bool mono_zero39(int pat, int k) {
  const bool zx = pat & 1, zy = pat & 2, zz = pat & 4;
  return k == 0 ? false : (k == 1 || k == 4 || k == 10) ? zx : (k == 2 || k == 5 || k == 11) ? zy
       : (k == 3 || k == 6 || k == 12) ? zz : k == 7 ? (zx || zy) : k == 8 ? (zx || zz)
       : (zy || zz);
}

// equilibrium of the directions of zero pattern kPat and their Jacobian terms
template <bool kShifted, int kPat, class Sink>
__host__ __device__ __forceinline__ void jac_group39(const double* __restrict__ lm, double rho,
                                                     double* __restrict__ Jd, Sink& sink) {
  constexpr int sh = kShifted ? 1 : 0;
  constexpr int cnt = kPatC.count[sh][kPat];
  constexpr int nd = kJDup39.n[sh];
#pragma unroll
  for (int j = 0; j < cnt; ++j) {
    const int i = kPatC.dir[sh][kPat][j];
    double ex = lm[0];
#pragma unroll
    for (int k = 1; k < kM39; ++k)
      if (!mono_zero39(kPat, k)) ex = fma((double)mono39<kShifted>(i, k), lm[k], ex);
    const double e = rho * exp(ex);
    sink(i, e);
    double me[kM39];
    me[0] = e;
#pragma unroll
    for (int k = 1; k < kM39; ++k) me[k] = mono_zero39(kPat, k) ? 0.0 : (double)mono39<kShifted>(i, k) * e;
#pragma unroll
    for (int p = 0; p < nd; ++p) {
      const int a = kJDup39.rl[sh][p], b = kJDup39.rk[sh][p];
      if (!mono_zero39(kPat, a) && !mono_zero39(kPat, b))
        Jd[p] = fma((double)mono39<kShifted>(i, a), me[b], Jd[p]);
    }
  }
}

// Newton on the moment constraints; lm in/out, lm_prev = multipliers of the
// last iteration before its update; sink(i, e) receives each iteration's
// equilibrium (so the caller keeps the last one).  Returns iterations, or -1
// at the cap.
template <bool kShifted, class Sink = NoSink39>
__host__ __device__ inline int lm_newton39(double rho, const double tg[kM39], double lm[kM39],
                                           double lm_prev[kM39], int maxit, Sink sink = Sink()) {
  double mx = 1.0;
  int it = 0;
  while (mx > 1.0e-7) {
    if (it >= maxit) return -1;
    ++it;
    constexpr int nd = kJDup39.n[kShifted ? 1 : 0];
    double Jd[nd];
#pragma unroll
    for (int p = 0; p < nd; ++p) Jd[p] = 0.0;
    jac_group39<kShifted, 0>(lm, rho, Jd, sink);
    jac_group39<kShifted, 1>(lm, rho, Jd, sink);
    jac_group39<kShifted, 2>(lm, rho, Jd, sink);
    jac_group39<kShifted, 3>(lm, rho, Jd, sink);
    jac_group39<kShifted, 4>(lm, rho, Jd, sink);
    jac_group39<kShifted, 5>(lm, rho, Jd, sink);
    jac_group39<kShifted, 6>(lm, rho, Jd, sink);
    jac_group39<kShifted, 7>(lm, rho, Jd, sink);
    double J[kJ39];
#pragma unroll
    for (int q = 0; q < kJ39; ++q) J[q] = Jd[kJDup39.cls[kShifted ? 1 : 0][q]];
    // residual: current moments (row 0 of J) minus targets
    double r[kM39];
#pragma unroll
    for (int k = 0; k < kM39; ++k) r[k] = J[jidx(k, 0)] - tg[k];
    // L D L^T in place: J[jidx(k,k)] <- D_k, J[jidx(k,l)] <- L_kl (l < k);
    // one division per pivot, the rest multiply by its reciprocal
    // v[m] = L_km D_m of the current row: one FMA per update term
    double rD[kM39];
#pragma unroll
    for (int k = 0; k < kM39; ++k) {
      double v[kM39];
#pragma unroll
      for (int l = 0; l < k; ++l) {
        double s = J[jidx(k, l)];
#pragma unroll
        for (int m = 0; m < l; ++m) s = fma(-v[m], J[jidx(l, m)], s);
        v[l] = s;
        J[jidx(k, l)] = s * rD[l];
      }
      double d = J[jidx(k, k)];
#pragma unroll
      for (int m = 0; m < k; ++m) d = fma(-v[m], J[jidx(k, m)], d);
      J[jidx(k, k)] = d;
      rD[k] = 1.0 / d;
    }
    // solve L y = r, D z = y, L^T delta = z
#pragma unroll
    for (int k = 0; k < kM39; ++k) {
#pragma unroll
      for (int m = 0; m < k; ++m) r[k] -= J[jidx(k, m)] * r[m];
    }
#pragma unroll
    for (int k = 0; k < kM39; ++k) r[k] *= rD[k];
#pragma unroll
    for (int k = kM39 - 1; k >= 0; --k) {
#pragma unroll
      for (int m = k + 1; m < kM39; ++m) r[k] -= J[jidx(m, k)] * r[m];
    }
    mx = 0.0;
#pragma unroll
    for (int k = 0; k < kM39; ++k) {
      lm_prev[k] = lm[k];
      lm[k] -= r[k];
      mx = fmax(mx, fabs(r[k]));
    }
    if (!(mx == mx)) return -1;
  }
  return it;
}

// exp(-s) for the small argument left by the last Newton update (|s| below
// 1e-3: five Taylor terms, error below 2e-21), else exp
__host__ __device__ inline double expm_small(double s) {
  if (fabs(s) < 1e-3) return 1.0 - s * (1.0 - s * (0.5 - s * (1.0 / 6.0 - s * (1.0 / 24.0 - s / 120.0))));
  return exp(-s);
}

// quick_calcs.f95:24-58 (non-dimensional viscosity; dim_* the SI values)
__host__ __device__ inline double sutherland39(double T, const Gas39& g, double* dim_mu,
                                               double* dim_temp) {
  const double const_s = (double)110.15f, temp_ref = (double)273.15f, visc_ref = 1.716e-5;
  const double unit_temp = T * g.t_ref * g.gama;
  const double unit_mu =
      visc_ref * pow(unit_temp / temp_ref, 1.5) * ((temp_ref + const_s) / (unit_temp + const_s));
  *dim_mu = unit_mu;
  *dim_temp = unit_temp;
  return unit_mu / g.mu_ref;
}

__host__ __device__ inline double kfit39(double ak, double bk, double ck, double dk, double ek,
                                         double unit_temp) {
  const double kek = log(unit_temp / 10000.0);
  const double k2 = kek * kek, k3 = k2 * kek, k4 = k3 * kek;
  return exp(ak * k4 + bk * k3 + ck * k2 + dk * kek + ek);
}

// quick_calcs.f95:70-215 with kappa_bottom = kappa_ref; false where the
// reference calls error_out
__host__ __device__ inline bool therm_cond39(double rho, double T, const Gas39& g, double* kappa) {
  const double kb = g.kappa_ref;
  const double unit_temp = T * g.t_ref * g.gama;
  const double unit_rho = rho * g.rho_ref;
  double unit_p = unit_temp * unit_rho * g.R;
  if (unit_temp <= 500.0) {
    *kappa = 418.1 * (5.9776e-6 * pow(unit_temp, 1.5) / (unit_temp + (double)194.4f)) / kb;
    return true;
  }
  if (unit_temp <= 2250.0) {
    unit_p = unit_rho * g.R * unit_temp / 101300.0;
    if (unit_p < 10.0 && unit_p > 0.1) { *kappa = kfit39(0.334316, 3.28202, 11.9939, 20.0944, 4.62882, unit_temp) / kb; return true; }
    if (unit_p > 10.0 && unit_p < 100.0) { *kappa = kfit39(0.413573, 3.83393, 13.1885, 20.7305, 4.27728, unit_temp) / kb; return true; }
    if (unit_p > 100.0 && unit_p < 1000.0) { *kappa = kfit39(0.208749, 1.92122, 6.58813, 10.7630, -1.27699, unit_temp) / kb; return true; }
    return false;
  }
  if (unit_temp <= 4250.0) {
    if (unit_p < 10.0 && unit_p > 0.1) { *kappa = kfit39(10.9992, 387.106, 387.282, 5.48304, -1.20106, unit_temp) / kb; return true; }
    if (unit_p > 10.0 && unit_p < 100.0) { *kappa = kfit39(82.1184, 308.927, 423.174, 250.668, 47.5889, unit_temp) / kb; return true; }
    if (unit_p > 100.0 && unit_p < 1000.0) { *kappa = kfit39(38.8677, 123.284, 144.224, 72.8083, 0.684807, unit_temp) / kb; return true; }
    return false;
  }
  return false;
}

}  // namespace uclbs
