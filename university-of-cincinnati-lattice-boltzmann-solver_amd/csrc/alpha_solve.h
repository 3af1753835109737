// Entropic over-relaxation parameter alpha: device-side root-finder.
//
// alpha is the nontrivial root of the entropy equality H(f + alpha*D) = H(f),
// D = feq - f, following the Newton definition of the reference
// (looping/timestepping/calculate_alpha.f95:416-556, commented out there):
// positivity bracket alpha_max = min_{D_i<0} -f_i/D_i (:424-445), the rule
// "alpha_max < 2 => alpha = alpha_max" (:449-455) and the clamp to alpha_max
// (:529-536).  Two changes make it safe on a GPU: a bracket with bisection
// fallback and an iteration cap, and the entropy function is evaluated in the
// conditioned form
//     g(a)  = sum_i f_i [(1 + a x_i) ln(1 + a x_i) - a x_i - a x_i ln(1 + x_i)],
//     g'(a) = sum_i D_i [ln(1 + a x_i) - ln(1 + x_i)],     x_i = D_i / f_i,
// which equals H(f + aD) - H(f) when D conserves mass and momentum and
// ln(feq_i / w_i) is affine in c_i (true for the D3Q27 product-form
// equilibrium).  The term -a x_i sums to -a sum(D) = 0 in exact arithmetic;
// keeping it inside phi(y) = (1+y) ln(1+y) - y stops the FP64 rounding of the
// conserved moments of feq from biasing the root.  For max|x_i| small the same g is evaluated from the power
// sums S_n = sum_i f_i x_i^n:
//     g(a)  = sum_{n>=2} (-1)^n S_n (a^n/(n(n-1)) - a/(n-1)),
//     g'(a) = sum_{n>=2} (-1)^n S_n (a^(n-1) - 1)/(n-1),
// truncated at an order whose remainder is below FP64 round-off for the
// node's max|x_i|; each Newton iteration then costs O(order) scalar FP64 work
// instead of 27 logarithms.  All of it runs in FP64.
#pragma once
#include <hip/hip_runtime.h>

namespace uclbs {

enum : int { ALPHA_NEWTON = 0, ALPHA_TRIVIAL = 1, ALPHA_BOUND = 2, ALPHA_NONPOS = 3,
             ALPHA_NOCONV = 4, ALPHA_DEFER = 5 /* left to the far-from-equilibrium kernel */ };

struct AlphaOut {
  double alpha;
  int iters;
  int status;
  bool logpath = false;  // g needed direct logarithms (max|x| beyond the series range)
  bool fallback = false; // the learned start failed and the plain start (alpha = 2) ran
};

// below this max|x| the node is at equilibrium and alpha = 2 (the "fi = fieq"
// branch, calculate_alpha.f95:430-431 and :551-556)
constexpr double kTrivialX = 1e-12;

// Largest max|x| for which the order-N truncation of the power-sum series is
// used.  The remainder of g relative to S2 is about (2.2 m)^(N-1) / (N(N-1))
// for Newton iterates a <= 2.2; the host sets the bound from a target eps
// (1e-16 in FP64 mode, i.e. exact to round-off; tol / 2 in FP32 mode, so
// truncation adds at most half the Newton tolerance to the residual of g).
struct SeriesThr {
  double m5, m7, m10, m16;  // general path orders
  float m_fast;             // order kFastOrder of the fused FP32-storage pass
  float m_mid;              // order kMidOrder, the second in-register tier
};
constexpr int kFastOrder = 6;
constexpr int kMidOrder = 12;
inline double series_limit(int N, double eps) {
  return __builtin_pow(eps * N * (N - 1.0), 1.0 / (N - 1.0)) / 2.2;
}
inline SeriesThr make_series_thr(double eps) {
  return SeriesThr{series_limit(5, eps), series_limit(7, eps), series_limit(10, eps),
                   series_limit(16, eps), (float)series_limit(kFastOrder, eps),
                   (float)series_limit(kMidOrder, eps)};
}

// ln(1 + x) for x > -1 in FP64, about 35 VALU operations: u = 1 + x = 2^e m
// with m in [sqrt(1/2), sqrt(2)), ln m = 2 atanh(s), s = (m-1)/(m+1), |s| <=
// 0.1716, summed to s^23 (remainder < 1e-18); ln 2 split in a high part exact
// in e*ln2_hi and a low part; the rounding error of 1 + x is removed to first
// order, ln(1+x) = ln(u) - ((u - 1) - x)/u.
__host__ __device__ __forceinline__ double log1p_fast(double x) {
  const double u = 1.0 + x;
  const double c = (u - 1.0) - x;
  int e;
  double m = __builtin_frexp(u, &e);  // u = m 2^e, m in [0.5, 1)
  if (m < 0.70710678118654752440) {
    m *= 2.0;
    e -= 1;
  }
  const double s = (m - 1.0) / (m + 1.0);
  const double s2 = s * s;
  // 2 * sum_{k=0..11} s2^k / (2k+1)
  double p = 2.0 / 23.0;
  p = p * s2 + 2.0 / 21.0;
  p = p * s2 + 2.0 / 19.0;
  p = p * s2 + 2.0 / 17.0;
  p = p * s2 + 2.0 / 15.0;
  p = p * s2 + 2.0 / 13.0;
  p = p * s2 + 2.0 / 11.0;
  p = p * s2 + 2.0 / 9.0;
  p = p * s2 + 2.0 / 7.0;
  p = p * s2 + 2.0 / 5.0;
  p = p * s2 + 2.0 / 3.0;
  p = p * s2 + 2.0;
  const double ln2_hi = 6.93147180369123816490e-01;  // 0x3FE62E42FEE00000, 32 significant bits
  const double ln2_lo = 1.90821492927058770002e-10;
  const double de = (double)e;
  return de * ln2_hi + (de * ln2_lo + s * p) - c / u;
}

// 1/v for a normal v: the hardware estimate and two Newton steps (each at
// least doubles the correct bits), within an ulp or so of the quotient
__device__ __forceinline__ double drcp(double v) {
  double r = __builtin_amdgcn_rcp(v);
  r = fma(fma(-v, r, 1.0), r, r);
  r = fma(fma(-v, r, 1.0), r, r);
  return r;
}

// log1p_fast with both divisions replaced by reciprocals (device only): the
// quotient s only needs relative accuracy (its error enters ln as 2 ds with
// |s| <= 0.1716), and c/u is a first-order correction of size eps
__device__ __forceinline__ double log1p_dev(double x) {
  const double u = 1.0 + x;
  const double c = (u - 1.0) - x;
  int e;
  double m = __builtin_frexp(u, &e);
  if (m < 0.70710678118654752440) {
    m *= 2.0;
    e -= 1;
  }
  const double s = (m - 1.0) * drcp(m + 1.0);
  const double s2 = s * s;
  double p = 2.0 / 23.0;
  p = p * s2 + 2.0 / 21.0;
  p = p * s2 + 2.0 / 19.0;
  p = p * s2 + 2.0 / 17.0;
  p = p * s2 + 2.0 / 15.0;
  p = p * s2 + 2.0 / 13.0;
  p = p * s2 + 2.0 / 11.0;
  p = p * s2 + 2.0 / 9.0;
  p = p * s2 + 2.0 / 7.0;
  p = p * s2 + 2.0 / 5.0;
  p = p * s2 + 2.0 / 3.0;
  p = p * s2 + 2.0;
  const double ln2_hi = 6.93147180369123816490e-01;
  const double ln2_lo = 1.90821492927058770002e-10;
  const double de = (double)e;
  return de * ln2_hi + (de * ln2_lo + s * p) - c * __builtin_amdgcn_rcp(u);
}

// ln(1 + y) for y > -1 through a 64-entry table over the mantissa: u = 1 + y
// = m 2^e with m in [1, 2), c_k = 1 + (k + 1/2)/64 the centre of m's 1/64
// interval, inv_k = 1/c_k, r = m inv_k - 1 (|r| < 1/128, one FMA) and
// ln u = e ln 2 + L_k + ln(1 + r), L_k = -ln(inv_k), ln(1 + r) by its series
// to r^7 (|r|^8 / 8 < 2e-18).  The rounding of 1 + y is not corrected, so
// the absolute error is up to 1.2e-16: in g = sum f ((1+y) ln(1+y) - y) it
// stays below the acceptance |g| <= tol S2 (tol 1e-10) while S2 > 1e-5 rho,
// as on the far nodes this serves; a node below would report no convergence,
// not a wrong root.  14 FP64 operations and two table reads where log1p_dev
// takes about 30.
struct LogTable {
  double inv[64], L[64];
};
// entry k of the table (one lane per entry)
__device__ __forceinline__ void log_table_entry(LogTable& t, int k) {
  const double inv = 1.0 / (1.0 + (k + 0.5) / 64.0);
  t.inv[k] = inv;
  t.L[k] = -log1p_dev(inv - 1.0);  // inv in (0.5, 1): inv - 1 is exact
}
__device__ __forceinline__ double log1p_tab(double y, const LogTable& t) {
  const double u = 1.0 + y;
  int e;
  const double m = 2.0 * __builtin_frexp(u, &e);
  const int k = (int)((m - 1.0) * 64.0);
  const double r = fma(m, t.inv[k], -1.0);
  double p = 1.0 / 7.0;
  p = fma(p, r, -1.0 / 6.0);
  p = fma(p, r, 1.0 / 5.0);
  p = fma(p, r, -0.25);
  p = fma(p, r, 1.0 / 3.0);
  p = fma(p, r, -0.5);
  p = fma(p, r, 1.0);
  const double ln2_hi = 6.93147180369123816490e-01;
  const double ln2_lo = 1.90821492927058770002e-10;
  const double de = (double)(e - 1);
  return (de * ln2_hi + t.L[k]) + (de * ln2_lo + p * r);
}

// phi(y) = (1 + y) ln(1 + y) - y, series for small |y|, direct otherwise
__device__ __forceinline__ double phi1p(double y) {
  if (fabs(y) < 0.1) {
    double s = 0.0, p = y * y;
#pragma unroll
    for (int n = 2; n <= 20; ++n) {
      s += ((n & 1) ? -p : p) * (1.0 / (n * (n - 1.0)));
      p *= y;
    }
    return s;
  }
  return (1.0 + y) * log1p(y) - y;
}

template <int N>
__device__ __forceinline__ void series_g(const double* S, double a, double& g, double& gp) {
  double an1 = a, sg = 0.0, sgp = 0.0;
#pragma unroll
  for (int n = 2; n <= N; ++n) {
    const double an = an1 * a;
    const double c1 = 1.0 / (n * (n - 1.0)), c2 = 1.0 / (n - 1.0);
    const double tg = S[n] * (an * c1 - a * c2);
    const double tgp = S[n] * ((an1 - 1.0) * c2);
    if (n & 1) { sg -= tg; sgp -= tgp; } else { sg += tg; sgp += tgp; }
    an1 = an;
  }
  g = sg;
  gp = sgp;
}

// Same series with the orders split over two arrays: S[2..N1] and Sh[0..N2-N1-1]
// holding orders N1+1..N2 (keeps the second tier's registers small).
template <int N1, int N2>
__device__ __forceinline__ void series_g2(const double* S, const double* Sh, double a, double& g,
                                          double& gp) {
  double an1 = a, sg = 0.0, sgp = 0.0;
#pragma unroll
  for (int n = 2; n <= N2; ++n) {
    const double an = an1 * a;
    const double c1 = 1.0 / (n * (n - 1.0)), c2 = 1.0 / (n - 1.0);
    const double Sn = n <= N1 ? S[n] : Sh[n - N1 - 1];
    const double tg = Sn * (an * c1 - a * c2);
    const double tgp = Sn * ((an1 - 1.0) * c2);
    if (n & 1) { sg -= tg; sgp -= tgp; } else { sg += tg; sgp += tgp; }
    an1 = an;
  }
  g = sg;
  gp = sgp;
}

// Safeguarded Newton shared by both evaluations of g.  Iterations count Newton
// updates taken before |g(a)| <= tol * S2 (S2 = sum f x^2 ~ g'(root), so the
// test bounds |a - root| by about tol).
template <class Eval>
__device__ __forceinline__ AlphaOut newton_alpha(Eval&& ev, double S2, double amax,
                                                 double alpha0, double tol, int maxit) {
  double lo = 1.0, hi = amax < 1e6 ? amax : 1e6;
  bool hi_is_bound = amax < 1e6;
  double a = (alpha0 > lo && alpha0 < hi) ? alpha0 : 2.0;
  AlphaOut r{2.0, 0, ALPHA_NEWTON, false};
  for (int k = 0;; ++k) {
    double g, gp;
    ev(a, g, gp);
    if (fabs(g) <= tol * S2) { r.alpha = a; r.iters = k; return r; }
    if (g < 0.0 && a >= amax) { r.alpha = amax; r.iters = k; r.status = ALPHA_BOUND; return r; }
    if (k >= maxit) { r.alpha = a; r.iters = k; r.status = ALPHA_NOCONV; return r; }
    if (g < 0.0) lo = a; else { hi = a; hi_is_bound = false; }
    double an = a - g / gp;
    if (!(an > lo && an < hi)) an = (hi_is_bound && an >= hi) ? amax : 0.5 * (lo + hi);
    if (an == amax) hi_is_bound = false;
    a = an;
  }
}

// Root of g for a node far from equilibrium, from sweeps of logarithms at a
// centre a0 and the exact re-expansion of g about it.  With d = a - a0,
// A_i = 1 + a0 x_i and z_i = x_i / A_i, 1 + a x_i = A_i (1 + d z_i) and
//   g(a0 + d) = g(a0) + g'(a0) d + sum_{n>=2} (-1)^n T_n d^n / (n(n-1)),
//   T_n = sum_i D_i z_i^(n-1),
// from (1+t) ln(1+t) = t + sum_{n>=2} (-1)^n t^n / (n(n-1)) at t = d z_i and
// f_i A_i z_i^n = D_i z_i^(n-1).  Since D_i z_i = f_i x_i^2 / A_i >= 0,
// |T_n| <= zmax^(n-2) T_2 and the orders beyond N add at most
//   R = T_2 d^2 (|d| zmax)^(N-1) / (N (N+1) (1 - |d| zmax)).
// One sweep (Q logarithms) gives g(a0), g'(a0) and T_2..T_N; Newton on the
// order-N polynomial (scalar work) gives d, and a0 + d is the root when the
// polynomial's residual plus R is within tol * S2.  Otherwise the next sweep
// is centred at a0 + d, kept inside the Newton bracket and within
// |d| zmax <= 1/2.  From a seed near the root one sweep suffices where Newton
// on direct logarithms evaluates g two to four times.  Bracket, bound and
// iteration cap follow newton_alpha; iters counts the updates of the centre
// or of the accepted root, as Newton counts its updates.
// dx(i, d, x) yields D_i and x_i in FP64, fx(i) yields f_i; K = sum D ln(1+x).
// kUnroll: unroll factor of the sweep over i (1 when dx / fx read memory).
// lg(y): ln(1 + y) (log1p_dev or log1p_tab).
template <int Q, int N, int kUnroll, class DX, class FX, class LG>
__device__ __forceinline__ AlphaOut recentred_alpha(DX&& dx, FX&& fx, LG&& lg, double K, double S2,
                                                    double amax, double alpha0, double tol,
                                                    int maxit) {
  double lo = 1.0, hi = amax < 1e6 ? amax : 1e6;
  bool hi_is_bound = amax < 1e6;
  double a = (alpha0 > lo && alpha0 < hi) ? alpha0 : 2.0;
  AlphaOut r{2.0, 0, ALPHA_NEWTON, true};
  for (int k = 0;; ++k) {
    double g = 0.0, gp = 0.0, zmax = 0.0;
    // the polynomial's coefficients only steer the step (g and g' stay FP64
    // and decide acceptance), so they are summed in FP32; their rounding
    // enters the acceptance bound below
    float Tf[N + 1];
#pragma unroll
    for (int n = 2; n <= N; ++n) Tf[n] = 0.0f;
    bool edge = false;  // a at the positivity bound: some 1 + a x_i <= 0
#pragma unroll kUnroll
    for (int i = 0; i < Q; ++i) {
      double d, x;
      dx(i, d, x);
      const double f = fx(i);
      const double y = a * x;
      const double A = 1.0 + y;
      if (A <= 0.0) { g -= f * y; edge = true; continue; }  // limit of (1+y) ln(1+y) at 0
      const double la = lg(y);
      g += f * (A * la - y);
      gp += d * la;
      const double z = x * drcp(A);
      zmax = fmax(zmax, fabs(z));
      const float zf = (float)z;
      float t = (float)d * zf;
      Tf[2] += t;
#pragma unroll
      for (int n = 3; n <= N; ++n) { t *= zf; Tf[n] += t; }
    }
    double T[N + 1];
#pragma unroll
    for (int n = 2; n <= N; ++n) T[n] = (double)Tf[n];
    g -= a * K;
    gp -= K;
    if (fabs(g) <= tol * S2) { r.alpha = a; r.iters = k; return r; }
    if (g < 0.0 && a >= amax) { r.alpha = amax; r.iters = k; r.status = ALPHA_BOUND; return r; }
    if (k >= maxit) { r.alpha = a; r.iters = k; r.status = ALPHA_NOCONV; return r; }
    if (g < 0.0) lo = a; else { hi = a; hi_is_bound = false; }
    if (edge) { a = 0.5 * (lo + hi); continue; }
    // Newton on the polynomial from d = 0, |d| within the trust radius
    const double rad = zmax > 0.0 ? 0.5 / zmax : 1e300;
    double dl = lo - a, dh = hi - a;
    double dd = 0.0, p = g;
    bool ok = true;
    for (int j = 0; j < 8; ++j) {
      double pd = gp, dn = dd;  // p'(d): gp + sum (-1)^n T_n d^(n-1) / (n-1)
#pragma unroll
      for (int n = 2; n <= N; ++n) {
        const double tn = T[n] * dn / (n - 1.0);
        pd += (n & 1) ? -tn : tn;
        dn *= dd;
      }
      double st = dd - p / pd;
      if (!(st > dl && st < dh) || fabs(st) > rad) { ok = false; dd = st; break; }
      dd = st;
      // p(d) = g + gp d + sum (-1)^n T_n d^n / (n(n-1))
      double pv = g + gp * dd, dn2 = dd * dd;
#pragma unroll
      for (int n = 2; n <= N; ++n) {
        const double tn = T[n] * dn2 / (n * (n - 1.0));
        pv += (n & 1) ? -tn : tn;
        dn2 *= dd;
      }
      p = pv;
      if (fabs(p) <= 0.25 * tol * S2) break;
    }
    const double q = fabs(dd) * zmax;
    double rem = 1e300;
    if (ok && q < 0.5) {
      double qp = 1.0;
#pragma unroll
      for (int n = 0; n < N - 1; ++n) qp *= q;
      rem = T[2] * dd * dd * qp / (N * (N + 1.0) * (1.0 - q));
      // the FP32 coefficients: relative error below 1e-6 (27 terms summed),
      // times the size of each polynomial term at dd
      double dn = dd * dd, terms = 0.0;
#pragma unroll
      for (int n = 2; n <= N; ++n) {
        terms += fabs(T[n]) * dn / (n * (n - 1.0));
        dn *= fabs(dd);
      }
      rem += 1e-6 * terms;
    }
    if (ok && fabs(p) + rem <= tol * S2) { r.alpha = a + dd; r.iters = k + 1; return r; }
    // next centre: the polynomial's step where it stays in the bracket and the
    // trust radius, else the bound or the middle of the bracket
    double an = a + (fabs(dd) > rad ? (dd > 0.0 ? rad : -rad) : dd);
    if (!(an > lo && an < hi)) an = (hi_is_bound && an >= hi) ? amax : 0.5 * (lo + hi);
    if (an == amax) hi_is_bound = false;
    a = an;
  }
}

template <int N, class XF>
__device__ __forceinline__ void power_sums(XF&& dx, int q, double* S) {
#pragma unroll
  for (int n = 2; n <= N; ++n) S[n] = 0.0;
#pragma unroll
  for (int i = 0; i < q; ++i) {
    double d, x;
    dx(i, d, x);
    double t = d * x;  // f x^2
    S[2] += t;
#pragma unroll
    for (int n = 3; n <= N; ++n) { t *= x; S[n] += t; }
  }
}

// dx(i, d, x) yields D_i and x_i in FP64; fx(i) yields f_i in FP64.
template <int Q, class DX, class FX>
__device__ AlphaOut solve_alpha(DX&& dx, FX&& fx, double m, double xmin, double S2,
                                double alpha0, double tol, int maxit, const SeriesThr& thr) {
  const double amax = xmin < 0.0 ? -1.0 / xmin : 1e300;
  if (m < kTrivialX) return AlphaOut{2.0, 0, ALPHA_TRIVIAL, false};
  if (amax < 2.0) return AlphaOut{amax, 0, ALPHA_BOUND, false};
  double S[17];
  if (m <= thr.m7) {
    if (m <= thr.m5) {
      power_sums<5>(dx, Q, S);
      return newton_alpha([&](double a, double& g, double& gp) __attribute__((always_inline)) { series_g<5>(S, a, g, gp); },
                          S2, amax, alpha0, tol, maxit);
    }
    power_sums<7>(dx, Q, S);
    return newton_alpha([&](double a, double& g, double& gp) __attribute__((always_inline)) { series_g<7>(S, a, g, gp); }, S2,
                        amax, alpha0, tol, maxit);
  }
  if (m <= thr.m10) {
    power_sums<10>(dx, Q, S);
    return newton_alpha([&](double a, double& g, double& gp) __attribute__((always_inline)) { series_g<10>(S, a, g, gp); },
                        S2, amax, alpha0, tol, maxit);
  }
  if (m <= thr.m16) {
    power_sums<16>(dx, Q, S);
    return newton_alpha([&](double a, double& g, double& gp) __attribute__((always_inline)) { series_g<16>(S, a, g, gp); },
                        S2, amax, alpha0, tol, maxit);
  }
  // large non-equilibrium: direct logarithms, x_i and ln(1 + x_i) formed once
  // (with log1p_fast each evaluation costs 27 logarithms of ~35 operations)
  double xs[Q], l1[Q];
#pragma unroll
  for (int i = 0; i < Q; ++i) {
    double d;
    dx(i, d, xs[i]);
    l1[i] = log1p_fast(xs[i]);
  }
  AlphaOut r = newton_alpha(
      [&](double a, double& g, double& gp) __attribute__((always_inline)) {
        double sg = 0.0, sgp = 0.0;
        bool inf_gp = false;
#pragma unroll
        for (int i = 0; i < Q; ++i) {
          const double x = xs[i];
          const double f = fx(i);
          const double y = a * x;
          if (1.0 + y <= 0.0) { sg += f * (-y - y * l1[i]); inf_gp = true; continue; }
          const double la = log1p_fast(y);
          // (1+y) ln(1+y) - y loses only absolute accuracy O(eps*|y|), far
          // below tol * S2
          sg += f * ((1.0 + y) * la - y - y * l1[i]);
          sgp += f * x * (la - l1[i]);
        }
        g = sg;
        gp = inf_gp ? __builtin_huge_val() : sgp;
      },
      S2, amax, alpha0, tol, maxit);
  r.logpath = true;
  return r;
}

}  // namespace uclbs
