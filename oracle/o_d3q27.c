/* D3Q27 entropic LBM oracle, FP64.  TEST INFRASTRUCTURE ONLY (see uclbs_oracle.h).
 *
 * The reference has no D3Q27 collision (general/background_calcs/
 * directions_and_weights.f95:226-237 selects dir=27 only for mach <= 1e-4 and
 * looping/timestepping/collision.f95:211-222 handles only 19 and 39), so this
 * file states the published method the GPU library uses in its place, reusing the
 * reference's building blocks:
 *   - equilibrium: product-form entropic equilibrium (Ansumali, Karlin,
 *     Ottinger, Europhys. Lett. 63 (2003) 798), exact mass and momentum;
 *   - relaxation: fout = f + alpha*beta*(feq - f) with beta = 1/(6 nu + 1),
 *     the single-population form of collision.f95:289 with q_bar = 0;
 *   - alpha: Newton on the entropy equality H(f + a D) = H(f), the definition
 *     of calculate_alpha.f95:416-556 (bracket :424-445, alpha_max < 2 rule
 *     :449-455), made safe with a bisection fallback;
 *   - streaming: push f(x + c_i) = fout(x) (streaming.f95:365-368), done here
 *     as the equivalent pull on a periodic grid.
 *
 * Conditioned entropy function.  With x_i = D_i / f_i and sum_i D_i = 0,
 * sum_i D_i c_i = 0 and ln(feq_i/w_i) affine in c_i (true for the product
 * form), H(f + aD) - H(f) equals
 *     g(a) = sum_i f_i [ (1 + a x_i) ln(1 + a x_i) - a x_i - a x_i ln(1 + x_i) ],
 * which is O(x^2) per term and free of the cancellation of the direct form.
 *     g'(a) = sum_i D_i [ ln(1 + a x_i) - ln(1 + x_i) ].
 * g is convex, g(0) = 0, g'(1) = 0, so the wanted root lies in (1, a_max).
 */
#include "uclbs_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* 1-D factors of the product-form equilibrium: p[0] for c=-1, p[1] c=0, p[2] c=+1 */
static void psi_axis(double u, double p[3]) {
  double s = sqrt(1.0 + 3.0 * u * u);
  double b = 2.0 - s;
  p[1] = b;
  p[2] = b * (2.0 * u + s) / (1.0 - u);
  p[0] = b * (1.0 - u) / (2.0 * u + s);
}

void o27_feq(double rho, double ux, double uy, double uz, double feq[27]) {
  double px[3], py[3], pz[3];
  psi_axis(ux, px);
  psi_axis(uy, py);
  psi_axis(uz, pz);
  for (int i = 0; i < 27; i++)
    feq[i] = rho * O27_W[i] * px[O27_CX[i] + 1] * py[O27_CY[i] + 1] * pz[O27_CZ[i] + 1];
}

/* calculate_flow_variables.f95:201-207: rho = sum f, rho u = sum c f */
void o27_moments(const double f[27], double *rho, double u[3]) {
  double r = 0, jx = 0, jy = 0, jz = 0;
  for (int i = 0; i < 27; i++) {
    r += f[i];
    jx += O27_CX[i] * f[i];
    jy += O27_CY[i] * f[i];
    jz += O27_CZ[i] * f[i];
  }
  *rho = r;
  u[0] = jx / r;
  u[1] = jy / r;
  u[2] = jz / r;
}

/* phi(y) = (1+y) ln(1+y) - y, accurate for small |y| (series) and large (direct). */
static double phi(double y) {
  if (fabs(y) < 0.1) {
    double s = 0.0, p = y * y;
    for (int n = 2; n <= 20; n++) {
      s += ((n & 1) ? -p : p) / (n * (n - 1.0));
      p *= y;
    }
    return s;
  }
  return (1.0 + y) * log1p(y) - y;
}

/* g(a) = sum f [phi(a x) - a x ln(1+x)]: the linear term a*sum(D), zero in exact
 * arithmetic, is left out explicitly so FP64 rounding of the conserved moments
 * of feq cannot bias the root. */
void o27_entropy_g(const double f[27], const double feq[27], double a, double *g, double *gp) {
  double sg = 0.0, sgp = 0.0;
  for (int i = 0; i < 27; i++) {
    double d = feq[i] - f[i];
    double x = d / f[i];
    double l1 = log1p(x);
    double y = a * x;
    if (1.0 + y <= 0.0) { /* at the positivity bound: (1+y) ln(1+y) -> 0, ln -> -inf */
      sg += f[i] * (-y - y * l1);
      sgp = INFINITY;
      continue;
    }
    double la = log1p(y);
    sg += f[i] * (phi(y) - y * l1);
    if (sgp != INFINITY) sgp += d * (la - l1);
  }
  *g = sg;
  *gp = sgp;
}

long double o27_H_ld(const long double f[27]) {
  long double h = 0.0L;
  for (int i = 0; i < 27; i++) h += f[i] * logl(f[i] / (long double)O27_W[i]);
  return h;
}

/* power-sum form: g(a) = sum_{n=2..N} (-1)^n S_n (a^n/(n(n-1)) - a/(n-1)) */
static void series_g(const double *S, int N, double a, double *g, double *gp) {
  double sg = 0.0, sgp = 0.0, an1 = a; /* a^(n-1) */
  for (int n = 2; n <= N; n++) {
    double an = an1 * a;
    double sgn = (n & 1) ? -1.0 : 1.0;
    sg += sgn * S[n] * (an / (n * (n - 1.0)) - a / (n - 1.0));
    sgp += sgn * S[n] * (an1 - 1.0) / (n - 1.0);
    an1 = an;
  }
  *g = sg;
  *gp = sgp;
}

#define TRIVIAL_X 1e-12

static o_alpha_result alpha_common(const double f[27], const double feq[27], double alpha0,
                                   double tol, int maxit, int series_order) {
  o_alpha_result r = {2.0, 0, O_ALPHA_NEWTON, 0.0};
  double x[27], m = 0.0, amax = INFINITY, S[40];
  for (int i = 0; i < 27; i++) {
    if (!(f[i] > 0.0)) { r.status = O_ALPHA_NONPOS; return r; }
  }
  double S2 = 0.0;
  for (int i = 0; i < 27; i++) {
    double d = feq[i] - f[i];
    x[i] = d / f[i];
    double ax = fabs(x[i]);
    if (ax > m) m = ax;
    if (x[i] < 0.0) {
      double b = -1.0 / x[i];
      if (b < amax) amax = b;
    }
    S2 += d * x[i];
  }
  if (m < TRIVIAL_X) { r.status = O_ALPHA_TRIVIAL; return r; }
  if (amax < 2.0) { r.alpha = amax; r.status = O_ALPHA_BOUND; return r; }
  if (series_order > 0) {
    for (int n = 2; n <= series_order; n++) S[n] = 0.0;
    for (int i = 0; i < 27; i++) {
      double p = f[i] * x[i] * x[i];
      for (int n = 2; n <= series_order; n++) { S[n] += p; p *= x[i]; }
    }
  }
  /* Safeguarded Newton: bracket [lo, hi] with g(lo) < 0; hi starts at the
   * positivity bound (or 1e6 when no x_i < 0).  A step that leaves the
   * bracket bisects, except that the bound itself is tried once: if g is
   * still negative there the root lies outside the positive region and
   * alpha = a_max (the reference's bound rule, calculate_alpha.f95:529-536). */
  double lo = 1.0, hi = (amax < 1e6) ? amax : 1e6;
  int hi_is_bound = amax < 1e6;
  double a = (alpha0 > lo && alpha0 < hi) ? alpha0 : 2.0;
  for (int k = 0;; k++) {
    double g, gp;
    if (series_order > 0) series_g(S, series_order, a, &g, &gp);
    else o27_entropy_g(f, feq, a, &g, &gp);
    r.resid = fabs(g) / S2;
    if (r.resid <= tol) { r.alpha = a; r.iters = k; return r; }
    if (g < 0.0 && a >= amax) { r.alpha = amax; r.iters = k; r.status = O_ALPHA_BOUND; return r; }
    if (k == maxit) { r.alpha = a; r.iters = k; r.status = O_ALPHA_NOCONV; return r; }
    if (g < 0.0) lo = a; else { hi = a; hi_is_bound = 0; }
    double an = a - g / gp;
    if (!(an > lo && an < hi)) an = (hi_is_bound && an >= hi) ? amax : 0.5 * (lo + hi);
    if (an == amax) hi_is_bound = 0;
    a = an;
  }
}

o_alpha_result o27_alpha(const double f[27], const double feq[27], double alpha0, double tol,
                         int maxit) {
  return alpha_common(f, feq, alpha0, tol, maxit, 0);
}

o_alpha_result o27_alpha_series(const double f[27], const double feq[27], double alpha0,
                                double tol, int maxit, int order) {
  if (order < 2) order = 2;
  if (order > 39) order = 39;
  return alpha_common(f, feq, alpha0, tol, maxit, order);
}

o_alpha_result o27_collide_node(const double f[27], double beta, double alpha0, double tol,
                                int maxit, int fixed_alpha_mode, double fout[27], double *rho,
                                double u[3]) {
  double feq[27];
  o27_moments(f, rho, u);
  o27_feq(*rho, u[0], u[1], u[2], feq);
  o_alpha_result ar = {2.0, 0, O_ALPHA_TRIVIAL, 0.0};
  if (!fixed_alpha_mode) ar = o27_alpha(f, feq, alpha0, tol, maxit);
  double ab = ar.alpha * beta;
  for (int i = 0; i < 27; i++) fout[i] = f[i] + ab * (feq[i] - f[i]);
  return ar;
}

static inline int wrap(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }

void o27_step_periodic(int nx, int ny, int nz, const double *fin, double *fout,
                       const o27_params *p, double *alpha_out, o_stats *st) {
  const int64_t N = (int64_t)nx * ny * nz;
  int64_t it = 0, nb = 0, nn = 0, nt = 0, nc = 0;
#ifdef _OPENMP
  if (p->nthreads > 0) omp_set_num_threads(p->nthreads);
#pragma omp parallel for collapse(2) schedule(static) reduction(+ : it, nb, nn, nt, nc)
#endif
  for (int z = 0; z < nz; z++)
    for (int y = 0; y < ny; y++)
      for (int x = 0; x < nx; x++) {
        double f[27], fo[27], rho, u[3];
        int64_t node = x + (int64_t)nx * (y + (int64_t)ny * z);
        for (int i = 0; i < 27; i++) {
          int xs = wrap(x - O27_CX[i], nx), ys = wrap(y - O27_CY[i], ny),
              zs = wrap(z - O27_CZ[i], nz);
          f[i] = fin[(int64_t)i * N + xs + (int64_t)nx * (ys + (int64_t)ny * zs)];
        }
        o_alpha_result ar =
            o27_collide_node(f, p->beta, 0.0, p->tol, p->maxit, p->alpha_mode, fo, &rho, u);
        for (int i = 0; i < 27; i++) fout[(int64_t)i * N + node] = fo[i];
        if (alpha_out) alpha_out[node] = ar.alpha;
        it += ar.iters;
        nb += ar.status == O_ALPHA_BOUND;
        nn += ar.status == O_ALPHA_NONPOS;
        nt += ar.status == O_ALPHA_TRIVIAL;
        nc += ar.status == O_ALPHA_NOCONV;
      }
  if (st) {
    st->iters += it;
    st->nodes += N;
    st->n_bound += nb;
    st->n_nonpos += nn;
    st->n_trivial += nt;
    st->n_noconv += nc;
  }
}

void o27_init_tgv(int nx, int ny, int nz, double u0, double noise_amp, uint64_t seed,
                  double *f) {
  const int64_t N = (int64_t)nx * ny * nz;
  const double tp = 6.283185307179586476925286766559;
#ifdef _OPENMP
#pragma omp parallel for collapse(2) schedule(static)
#endif
  for (int z = 0; z < nz; z++)
    for (int y = 0; y < ny; y++)
      for (int x = 0; x < nx; x++) {
        int64_t node = x + (int64_t)nx * (y + (int64_t)ny * z);
        double X = tp * x / nx, Y = tp * y / ny, Z = tp * z / nz;
        double u[3];
        u[0] = u0 * sin(X) * cos(Y) * cos(Z);
        u[1] = -u0 * cos(X) * sin(Y) * cos(Z);
        u[2] = 0.0;
        for (int c = 0; c < 3; c++)
          u[c] += noise_amp * (2.0 * o_uniform((uint64_t)node, (uint32_t)c, seed) - 1.0);
        double feq[27];
        o27_feq(1.0, u[0], u[1], u[2], feq);
        for (int i = 0; i < 27; i++) f[(int64_t)i * N + node] = feq[i];
      }
}

void o27_macro(int n, const double *f, double *rho, double *ux, double *uy, double *uz) {
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
  for (int k = 0; k < n; k++) {
    double fl[27], r, u[3];
    for (int i = 0; i < 27; i++) fl[i] = f[(int64_t)i * n + k];
    o27_moments(fl, &r, u);
    rho[k] = r;
    ux[k] = u[0];
    uy[k] = u[1];
    uz[k] = u[2];
  }
}
