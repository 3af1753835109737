/* D3Q39 compressible entropic step, FP64.  TEST INFRASTRUCTURE ONLY (see uclbs_oracle.h).
 *
 * Restates the reference's two-population (f, g) D3Q39 ELBM for one periodic
 * box at level lvl:
 *   constants     general/background_calcs/basic_fluids_calcs.f95:17-86,
 *                 utilities/viscosity_and_timing.f95:29-73,
 *                 general/background_calcs/timing_calcs.f95:24-26,
 *                 general/modules/constants.f95:14-17
 *   lattice       general/background_calcs/directions_and_weights.f95:436-536
 *                 (shifted: c_x + 1 for primary flow +x when Mach >= 1.2)
 *   moments       looping/timestepping/calculate_flow_variables.f95:201-228
 *   LM Newton     lagrangian_multipliers/calculate_lagrangian_mults.f95:337-1098
 *                 with lagrangian_multipliers/lu_decomp.f95
 *   equilibrium   amr/modules/amr_processes.f95:1592-1747 (fieq_comp_39[_shifted])
 *   alpha         looping/timestepping/calculate_alpha.f95:118-289 (the live
 *                 closed-form estimate)
 *   collision     looping/timestepping/collision.f95:206-293, find_q_bar.f95:33-220,
 *                 find_little_q_bar.f95:23-52, amr_processes.f95:1751-1814,
 *                 utilities/modules/quick_calcs.f95:24-215
 *   repair        looping/timestepping/damp_it.f95:64-177
 * Literals the reference writes without a kind suffix are default (single)
 * REAL widened to REAL(8); they are reproduced as (double)<float literal>.
 */
#include "uclbs_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define NQ 39
#define NM 13

/* constants.f95:14-17 */
#define C_PI 3.1415926535
#define C_MOLEC_MASS_AIR ((double)28.97f)
#define C_AVOGADRO 6.022e23
#define C_K_BOLTZ 1.38064852e-23
#define C_KAPPA_CONVERT 418.1
#define C_PLUS_EPS 1.24e-16
#define C_VSMALL 1.0e-14

void o39_velocities(int shifted, int cx[39], int cy[39], int cz[39]) {
  for (int i = 0; i < NQ; i++) {
    cx[i] = O39_CX[i] + (shifted ? 1 : 0);  /* directions_and_weights.f95:477-481 */
    cy[i] = O39_CY[i];
    cz[i] = O39_CZ[i];
  }
}

/* moment row of direction i in the Lagrange-multiplier order
 * chi, zeta(3), pi_xx, pi_yy, pi_zz, pi_xy, pi_xz, pi_yz, lambda(3) */
static void mrow(int x, int y, int z, int m[NM]) {
  int c2 = x * x + y * y + z * z;
  m[0] = 1; m[1] = x; m[2] = y; m[3] = z;
  m[4] = x * x; m[5] = y * y; m[6] = z * z;
  m[7] = x * y; m[8] = x * z; m[9] = y * z;
  m[10] = x * c2; m[11] = y * c2; m[12] = z * c2;
}

void o39_consts_make(o39_consts *c, double gama, double temperature_k, double pressure_pa,
                     double mach, double dx_m, int startup, int shifted, int turbulent, int lvl) {
  memset(c, 0, sizeof(*c));
  c->gama = gama;
  c->shifted = shifted;
  c->turbulent = turbulent;
  c->lvl = lvl;
  c->startup = startup;
  /* basic_fluids_calcs.f95:17-27 */
  double R = 287.052874;
  c->R = R;
  c->rho_ref = pressure_pa / (R * temperature_k);
  c->t_ref = temperature_k;
  c->p_ref = pressure_pa;
  c->cv = 1.0 / (gama - 1.0);
  c->cp = gama / (gama - 1.0);
  c->lead_g = 2.0 * c->cv - 3.0;                 /* collision.f95:73 */
  double temperature = 1.0 / gama, rho_inf = 1.0;  /* :31-33 */
  double sos = sqrt(gama * R * c->t_ref);          /* :37-38 */
  c->sos = sos;
  c->mach = mach;
  c->molec_weight = C_MOLEC_MASS_AIR / (C_AVOGADRO * 1000);  /* :86 */
  /* viscosity_and_timing.f95:29-68 */
  double S = (double)110.4f, visc_ref = (double)1.716E-5f, temp_ref = (double)273.15f;
  double viscosity = visc_ref * pow(c->t_ref / temp_ref, 1.5) * ((temp_ref + S) / (c->t_ref + S));
  c->mu_ref = c->rho_ref * sos * mach * 1;
  c->kappa_ref = 1.0;  /* kappa_bottom = 1 for the start-up conductivity */
  double kappa_base = 0.0;
  o39_therm_cond(rho_inf, temperature, c, &kappa_base);
  c->kappa_ref = rho_inf * sos * sos * sos / temperature;
  c->init_beta1 = 0.5001;
  c->init_beta2 = 0.5001;
  double b1 = 1.0 / ((2.0 * viscosity) / (rho_inf * temperature) + 1.0);
  double b2 = 1.0 / ((2.0 * kappa_base / c->kappa_ref) / (rho_inf * c->cp * temperature) + 1.0);
  c->beta1_scale = (b1 - c->init_beta1) / (double)startup;
  c->beta2_scale = (b2 - c->init_beta2) / (double)startup;
  /* timing_calcs.f95:24-26, level spacing dx / 2^lvl */
  double dx = dx_m / ldexp(1.0, lvl);
  c->dt = dx / sos;
  c->eps = c->dt / 1.0;
  c->maxit = 50;
}

/* quick_calcs.f95:24-58 (inter_lvl = .false.) */
double o39_sutherland(double T, const o39_consts *c, double *dim_mu, double *dim_temp) {
  const double const_s = (double)110.15f, temp_ref = (double)273.15f, visc_ref = 1.716e-5;
  double unit_temp = T * c->t_ref * c->gama;
  double unit_mu = visc_ref * pow(unit_temp / temp_ref, 1.5) * ((temp_ref + const_s) / (unit_temp + const_s));
  if (dim_mu) *dim_mu = unit_mu;
  if (dim_temp) *dim_temp = unit_temp;
  return unit_mu / c->mu_ref;
}

static double kfit(double ak, double bk, double ck, double dk, double ek, double unit_temp) {
  /* integer powers by repeated multiplication, as the Fortran ** with an
   * integer exponent evaluates them */
  double kek = log(unit_temp / 10000.0);
  double k2 = kek * kek, k3 = k2 * kek, k4 = k3 * kek;
  return exp(ak * k4 + bk * k3 + ck * k2 + dk * kek + ek);
}

/* quick_calcs.f95:70-215; kappa_bottom = c->kappa_ref.  Returns 0, or -1
 * where the reference calls error_out. */
int o39_therm_cond(double rho, double T, const o39_consts *c, double *kappa) {
  double kb = c->kappa_ref;
  double unit_temp = T * c->t_ref * c->gama;
  double unit_rho = rho * c->rho_ref;
  double unit_p = unit_temp * unit_rho * c->R;  /* Pa: the 2250-4250 K bands compare this */
  if (unit_temp <= 500.0) {
    *kappa = C_KAPPA_CONVERT * (5.9776e-6 * pow(unit_temp, 1.5) / (unit_temp + (double)194.4f)) / kb;
    return 0;
  }
  if (unit_temp <= 2250.0) {
    unit_p = unit_rho * c->R * unit_temp / 101300.0;  /* atm (quick_calcs.f95:101-103) */
    if (unit_p < 10.0 && unit_p > 0.1) { *kappa = kfit(0.334316, 3.28202, 11.9939, 20.0944, 4.62882, unit_temp) / kb; return 0; }
    if (unit_p > 10.0 && unit_p < 100.0) { *kappa = kfit(0.413573, 3.83393, 13.1885, 20.7305, 4.27728, unit_temp) / kb; return 0; }
    if (unit_p > 100.0 && unit_p < 1000.0) { *kappa = kfit(0.208749, 1.92122, 6.58813, 10.7630, -1.27699, unit_temp) / kb; return 0; }
    return -1;  /* the 0.01..0.1 atm band can never be selected (quick_calcs.f95:139) */
  }
  if (unit_temp <= 4250.0) {
    if (unit_p < 10.0 && unit_p > 0.1) { *kappa = kfit(10.9992, 387.106, 387.282, 5.48304, -1.20106, unit_temp) / kb; return 0; }
    if (unit_p > 10.0 && unit_p < 100.0) { *kappa = kfit(82.1184, 308.927, 423.174, 250.668, 47.5889, unit_temp) / kb; return 0; }
    if (unit_p > 100.0 && unit_p < 1000.0) { *kappa = kfit(38.8677, 123.284, 144.224, 72.8083, 0.684807, unit_temp) / kb; return 0; }
    return -1;
  }
  return -1;
}

/* fieq_comp_39_shifted (amr_processes.f95:1729-1747); the un-shifted
 * fieq_comp_39 (:1592-1723) is the same exponent written out per direction */
void o39_feq(const double lm[13], double rho, int shifted, double feq[39]) {
  int cx[NQ], cy[NQ], cz[NQ], m[NM];
  o39_velocities(shifted, cx, cy, cz);
  for (int i = 0; i < NQ; i++) {
    mrow(cx[i], cy[i], cz[i], m);
    double e = lm[0];
    for (int k = 1; k < NM; k++) e += (double)m[k] * lm[k];
    feq[i] = rho * exp(e);
  }
}

/* One Newton update through the LU factorisation of lu_decomp.f95: Doolittle
 * without pivoting (rows top-down, lower then upper), forward substitution
 * of J(1,k) - target_k, back substitution from the last unknown, lm -= delta.
 * ref_mode reproduces the two index slips of the reference,
 * lower(5,4) using lower(5,2)*upper(3,4) (:105) and z9 using lower(9,7)*z8
 * (:441); ref_mode = 0 is the exact factorisation.  Returns max|delta|. */
double o39_lu_step(const double *J, const double *mom, double *lm, int ref_mode) {
  /* mom = rho, u, v, w, Pxx, Pyy, Pzz, Pxy, Pxz, Pyz, qx, qy, qz; the
   * momentum targets are formed as rho * u (lu_decomp.f95:426-430) */
  double tgt[NM];
  for (int k = 0; k < NM; k++) tgt[k] = mom[k];
  for (int k = 1; k < 4; k++) tgt[k] = mom[0] * mom[k];
  double L[NM][NM], U[NM][NM], z[NM], d[NM];
  memset(L, 0, sizeof(L));
  memset(U, 0, sizeof(U));
  for (int j = 0; j < NM; j++) U[0][j] = J[j];
  for (int k = 1; k < NM; k++) {
    for (int j = 0; j < k; j++) {
      double s = J[k * NM + j];
      for (int m = 0; m < j; m++) {
        double lk = L[k][m];
        if (ref_mode && k == 4 && j == 3 && m == 2) lk = L[4][1];
        s -= lk * U[m][j];
      }
      L[k][j] = s / U[j][j];
    }
    for (int j = k; j < NM; j++) {
      double s = J[k * NM + j];
      for (int m = 0; m < k; m++) s -= L[k][m] * U[m][j];
      U[k][j] = s;
    }
  }
  for (int k = 0; k < NM; k++) {
    double s = J[k] - tgt[k];
    for (int m = 0; m < k; m++) {
      double zm = z[m];
      if (ref_mode && k == 8 && m == 6) zm = z[7];
      s -= L[k][m] * zm;
    }
    z[k] = s;
  }
  double mx = 0.0;
  for (int k = NM - 1; k >= 0; k--) {
    double s = z[k];
    for (int m = NM - 1; m > k; m--) s -= U[k][m] * d[m];
    d[k] = s / U[k][k];
  }
  for (int k = 0; k < NM; k++) {
    double a = fabs(d[k]);
    if (a > mx) mx = a;
    lm[k] -= d[k];
  }
  return mx;
}

/* Jacobian J(k,l) = sum_i m_k(c_i) m_l(c_i) expo_i (row-major) */
static void jacobian(const int *cx, const int *cy, const int *cz, const double *expo, double *J) {
  for (int k = 0; k < NM * NM; k++) J[k] = 0.0;
  for (int i = 0; i < NQ; i++) {
    int m[NM];
    mrow(cx[i], cy[i], cz[i], m);
    for (int k = 0; k < NM; k++)
      for (int l = 0; l < NM; l++) J[k * NM + l] += (double)(m[k] * m[l]) * expo[i];
  }
}

/* calculate_lagrangian_mults.f95:43-49, 337-1098: targets rho, rho u, P,
 * q = (rho |u|^2 + 5 rho T) u; Newton from lm until max|delta| <= 1e-7.
 * expo_last = the equilibrium of the last iteration's multipliers (before
 * its update: :1098).  Returns the iteration count, or -1 at the cap. */
int o39_lm_solve(double rho, const double u[3], double T, const double P[6], double lm[13],
                 double expo_last[39], int shifted, int ref_mode, int maxit) {
  int cx[NQ], cy[NQ], cz[NQ];
  o39_velocities(shifted, cx, cy, cz);
  double vm = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
  double qf = rho * vm + 5 * rho * T;
  double tgt[NM] = {rho, u[0], u[1], u[2], P[0], P[1], P[2], P[3], P[4], P[5],
                    qf * u[0], qf * u[1], qf * u[2]};
  double J[NM * NM];
  double max_diff = 1.0;
  int it = 0;
  while (max_diff > 1.0e-7) {
    if (it >= maxit) return -1;
    it++;
    o39_feq(lm, rho, shifted, expo_last);
    jacobian(cx, cy, cz, expo_last, J);
    max_diff = o39_lu_step(J, tgt, lm, ref_mode);
    if (!(max_diff == max_diff)) return -1;
  }
  return it;
}

/* calculate_alpha.f95:118-289 */
double o39_alpha_estimate(const double f[39], const double diff[39], double T, int64_t step,
                          const o39_consts *c, int *status) {
  const double alpha_hi = 2.125, alpha_start = 1.0e-6;
  double dim_mu, dim_temp;
  *status = 0;
  o39_sutherland(T, c, &dim_mu, &dim_temp);
  double mfp = dim_mu / c->p_ref * sqrt(C_PI * C_K_BOLTZ * dim_temp / (2 * c->molec_weight));
  double mft = mfp / sqrt(3 * C_K_BOLTZ * dim_temp / c->molec_weight);
  double beta_mfp = c->dt / (2 * mft + c->dt);
  double ft1 = 0, ft2 = 0, ft3 = 0, ft4 = 0, st1 = 0, st2 = 0, third = 0;
  double cs1 = 0, cs2 = 0, cs3 = 0;
  for (int i = 0; i < NQ; i++) {
    double df = f[i] < 0.0 ? fabs(f[i]) : f[i];
    double x = diff[i] / df;
    cs1 = cs1 + df * (2.0 * (x * x) / (2 + x));
    cs2 = cs2 + df * (x * x);
    if (x <= 0.0 && x >= -1.0) cs3 = cs3 + df * (x * x * x);
  }
  if (fabs(cs3) < C_PLUS_EPS) { *status = 1; return 2.0; }
  double asm_ = (cs2 - sqrt(cs2 * cs2 - 8.0 * cs3 * cs1)) / (2.0 * cs3);
  if (asm_ > alpha_hi) asm_ = alpha_hi - 1.0e-9;
  if (step < c->startup) asm_ = (asm_ - alpha_start) / (double)c->startup * (double)step;
  for (int i = 0; i < NQ; i++) {
    double df = f[i] < 0.0 ? fabs(f[i]) : f[i];
    double x1 = diff[i] / df, x2 = x1 * x1, x3 = x1 * x1 * x1;
    double f2 = x2 * df, f3 = x3 * df, f4 = (x2 * x2) * df, f5 = x2 * x3 * df, f6 = (x3 * x3) * df;
    if (x1 <= 0.0 && x1 >= -1.0) { ft1 += f3; ft2 += f4; ft3 += f5; ft4 += f6; }
    st1 += f2;
    if (x1 >= 0.0)
      st2 += f3 * (2.0 / (4 + asm_ * x1) + 1.0 / (4 + 2 * asm_ * x1) + 2.0 / (4 + 3 * asm_ * x1));
    third = third - (60 * f2 + 60 * f3 + 11 * f4) / (60 + 90 * x1 + 36 * x2 + 3 * x3);
  }
  double bm2 = beta_mfp * beta_mfp;
  double a2h = alpha_hi * alpha_hi, a3h = alpha_hi * alpha_hi * alpha_hi;
  double b3 = beta_mfp * beta_mfp * beta_mfp;
  double first = -bm2 * (ft1 / 6 - alpha_hi * beta_mfp * ft2 / 12 + a2h * bm2 * ft3 / 20 -
                         a3h * b3 * ft4 / 5);
  double second = st1 / 2 - 2 * asm_ * bm2 * st2 / 15;
  double na = 0.0, a;
  if (fabs(first) < C_VSMALL) {
    a = 2.0;
    *status = 1;
  } else {
    na = (-second + sqrt(second * second - 4 * first * third)) / (2 * first);
    if (na < asm_) { a = asm_; *status = 2; }
    else if (na > alpha_hi) { a = alpha_hi; *status = 3; }
    else a = na;
  }
  if (na != na) *status = -1;  /* error_out (:303-313) */
  return a;
}

/* find_q_bar.f95:52-220: the quasi-equilibrium correction of the
 * deviation df, from its central third moments Q_abc about u:
 * qbar_i = eps (sum_abc Q_abc c_a c_b c_c - 3 T sum_abc Q_abc delta_ab c_c) / (6 T^3)
 * (yi and kroen_val of directions_and_weights.f95:629-796) */
void o39_qbar(const double df[39], const double u[3], double T, double eps, int shifted,
              double qbar[39]) {
  int cx[NQ], cy[NQ], cz[NQ];
  o39_velocities(shifted, cx, cy, cz);
  double xxx = 0, yyy = 0, zzz = 0, xxy = 0, xxz = 0, xyy = 0, xzz = 0, yyz = 0, yzz = 0, xyz = 0;
  for (int i = 0; i < NQ; i++) {
    double a = cx[i] - u[0], b = cy[i] - u[1], e = cz[i] - u[2];
    xxx += a * a * a * df[i];
    yyy += b * b * b * df[i];
    zzz += e * e * e * df[i];
    xxy += a * a * b * df[i];
    xxz += a * a * e * df[i];
    xyy += b * b * a * df[i];
    xzz += e * e * a * df[i];
    yyz += b * b * e * df[i];
    yzz += e * e * b * df[i];
    xyz += a * b * e * df[i];
  }
  double qabc[27] = {xxx, xxy, xxz, xxy, xyy, xyz, xxz, xyz, xzz,
                     xxy, xyy, xyz, xyy, yyy, yyz, xyz, yyz, yzz,
                     xxz, xyz, xzz, xyz, yyz, yzz, xzz, yzz, zzz};
  double tc = 6.0 * T * T * T;
  for (int i = 0; i < NQ; i++) {
    int cv[3] = {cx[i], cy[i], cz[i]};
    double sy = 0.0, sk = 0.0;
    for (int a = 0; a < 3; a++)
      for (int b = 0; b < 3; b++)
        for (int e = 0; e < 3; e++) {
          double q = qabc[9 * a + 3 * b + e];
          sy += q * (double)(cv[a] * cv[b] * cv[e]);
          sk += q * (double)(a == b ? cv[e] : 0);   /* kroen_val: delta_ab c_c */
        }
    qbar[i] = eps * (sy - 3 * T * sk) / tc;
  }
}

/* find_little_q_bar.f95:23-52: hq_i = (eps / T) c_i . sum_j (c_j - u) dg_j */
void o39_little_qbar(const double dg[39], const double u[3], double T, double eps, int shifted,
                     double hq[39]) {
  int cx[NQ], cy[NQ], cz[NQ];
  o39_velocities(shifted, cx, cy, cz);
  double lq[3] = {0, 0, 0};
  for (int i = 0; i < NQ; i++) {
    lq[0] += (cx[i] - u[0]) * dg[i];
    lq[1] += (cy[i] - u[1]) * dg[i];
    lq[2] += (cz[i] - u[2]) * dg[i];
  }
  double ql = eps / T;
  for (int i = 0; i < NQ; i++) hq[i] = ql * (cx[i] * lq[0] + cy[i] * lq[1] + cz[i] * lq[2]);
}

/* collision.f95:206-293 for one node; feq from the stored multipliers */
int o39_collide(const double f[39], const double g[39], const double lm[13], double rho,
                const double u[3], double T, double alpha, int64_t step, const o39_consts *c,
                double fout[39], double gout[39]) {
  int cx[NQ], cy[NQ], cz[NQ];
  o39_velocities(c->shifted, cx, cy, cz);
  double feq[NQ], df[NQ], dg[NQ];
  o39_feq(lm, rho, c->shifted, feq);
  for (int i = 0; i < NQ; i++) {
    df[i] = feq[i] - f[i];
    dg[i] = c->lead_g * T * feq[i] - g[i];
  }
  /* epsilon_calculator (amr_processes.f95:1789-1814) */
  double eps = step < c->startup ? (((double)step + 1.0) / (double)c->startup) * c->eps : c->eps;
  double qbar[NQ], hq[NQ];
  o39_qbar(df, u, T, eps, c->shifted, qbar);
  o39_little_qbar(dg, u, T, eps, c->shifted, hq);
  /* quick_sutherlands / quick_therm_cond, beta_1/2_calculator */
  double mu = o39_sutherland(T, c, NULL, NULL), kappa = 0.0;
  int rc = o39_therm_cond(rho, T, c, &kappa);
  double b1, b2;
  if (step < c->startup) {
    b1 = c->init_beta1 + c->beta1_scale * (double)step;
    b2 = c->init_beta2 + c->beta2_scale * (double)step;
  } else {
    b1 = 1.0 / ((ldexp(1.0, c->lvl + 1) * mu) / (rho * T) + 1.0);
    b2 = 1.0 / ((2.0 * kappa) / (rho * c->cp * T) + 1.0);
  }
  for (int i = 0; i < NQ; i++) {
    fout[i] = f[i] + alpha * b1 * (df[i] - qbar[i]) + b2 * qbar[i];
    gout[i] = g[i] + alpha * b1 * (dg[i] - hq[i]) + b2 * hq[i];
  }
  return rc;
}

/* calculate_flow_variables.f95:201-228 (P_zz from the current w) */
void o39_moments(const double f[39], const double g[39], int shifted, double *rho, double u[3],
                 double *T, double P[6], double cv) {
  int cx[NQ], cy[NQ], cz[NQ];
  o39_velocities(shifted, cx, cy, cz);
  double r = 0, ux = 0, uy = 0, uz = 0, en = 0;
  for (int i = 0; i < NQ; i++) {
    double cm = (double)(cx[i] * cx[i] + cy[i] * cy[i] + cz[i] * cz[i]);
    r += f[i];
    ux += cx[i] * f[i];
    uy += cy[i] * f[i];
    en = en + cm * f[i] + g[i];
  }
  for (int i = 0; i < NQ; i++) uz += cz[i] * f[i];
  u[0] = ux / r; u[1] = uy / r; u[2] = uz / r;
  double vm = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
  *rho = r;
  *T = (en - r * vm) / (2 * cv * r);
  P[0] = r * (*T + u[0] * u[0]);
  P[1] = r * (*T + u[1] * u[1]);
  P[2] = r * (*T + u[2] * u[2]);
  P[3] = r * u[0] * u[1];
  P[4] = r * u[0] * u[2];
  P[5] = r * u[1] * u[2];
}

static inline int64_t wrap(int64_t a, int64_t n) { a %= n; return a < 0 ? a + n : a; }

/* Face codes of a box (x-, y-, z-, x+, y+, z+), the reference's edge-link
 * codes (amr/amr_code/new_node_values/state_directions_lvl.f95:94-420 and
 * shifted_state_directions_lvl.f95:194-686 assign them per face, in that
 * chain order): O39_FACE_PERIODIC wraps, -1000 freestream_edge_bc, -6 the
 * pure outflow of outlet type 6, -10 bounceback, -11 inviscid_edge. */
#define O39_FACE_PERIODIC 2147483647

static int axis_out(int64_t s, int64_t n, int f) { return f < 3 ? s < 0 : s >= n; }

/* Neighbour x +- c_i of damp_it.f95:64-123 on a box with faces: -1 when it
 * lies across a non-periodic face (a link whose code is negative, which the
 * reference repairs from the other side only), else its flat index. */
static int64_t damp_neighbour(int64_t a, int64_t b, int64_t cc, int dx, int dy, int dz, int nx,
                              int ny, int nz, const int32_t *fc, int ghost) {
  int64_t s[3] = {a + dx, b + dy, cc + dz};
  const int64_t n[3] = {nx, ny, nz};
  for (int d = 0; d < 3; d++) {
    if (s[d] < 0) {
      if (fc[d] != O39_FACE_PERIODIC) return -1;
      s[d] += n[d];
    } else if (s[d] >= n[d]) {
      if (fc[d + 3] != O39_FACE_PERIODIC) return -1;
      s[d] -= n[d];
    }
    /* a refined level's ghost layer holds no post-collision values */
    if (ghost > 0 && (s[d] < ghost || s[d] >= n[d] - ghost)) return -1;
  }
  return s[0] + n[0] * (s[1] + n[1] * s[2]);
}

/* damp_it.f95:64-123: a negative post-collision population is replaced by the
 * mean of its positive neighbours along +-c_i (or their mean with |f|, or
 * |f|); a neighbour across a non-periodic face counts as not positive, which
 * is the one-sided form of :109-123.  jacobi = 1 reads the neighbours' values
 * from before any repair (the order-free form the GPU uses); jacobi = 0 sweeps
 * in the reference order (direction, then z, y, x) and sees earlier repairs. */
static int in_ghost(int64_t a, int64_t b, int64_t cc, int nx, int ny, int nz, int ghost) {
  return ghost > 0 && (a < ghost || b < ghost || cc < ghost || a >= nx - ghost || b >= ny - ghost ||
                       cc >= nz - ghost);
}

static int64_t damp_core(int nx, int ny, int nz, const int32_t fc[6], int ghost, double *fo,
                         int shifted, int jacobi);

int64_t o39_damp_faces(int nx, int ny, int nz, const int32_t fc[6], double *fo, int shifted,
                       int jacobi) {
  return damp_core(nx, ny, nz, fc, 0, fo, shifted, jacobi);
}

/* a refined level's ghost layer (depth ghost) is not repaired: its nodes are
 * refilled from the coarser level */
static int64_t damp_core(int nx, int ny, int nz, const int32_t fc[6], int ghost, double *fo,
                         int shifted, int jacobi) {
  int cx[NQ], cy[NQ], cz[NQ];
  o39_velocities(shifted, cx, cy, cz);
  const int64_t n = (int64_t)nx * ny * nz;
  double *src = fo;
  if (jacobi) {
    src = (double *)malloc(sizeof(double) * (size_t)(n * NQ));
    memcpy(src, fo, sizeof(double) * (size_t)(n * NQ));
  }
  int64_t nrep = 0;
  for (int i = 0; i < NQ; i++) {
    for (int64_t k = 0; k < n; k++) {
      double v = src[(int64_t)i * n + k];
      if (!(v < 0.0)) continue;
      int64_t a = k % nx, b = (k / nx) % ny, cc = k / ((int64_t)nx * ny);
      if (in_ghost(a, b, cc, nx, ny, nz, ghost)) continue;
      int64_t kp = damp_neighbour(a, b, cc, cx[i], cy[i], cz[i], nx, ny, nz, fc, ghost);
      int64_t km = damp_neighbour(a, b, cc, -cx[i], -cy[i], -cz[i], nx, ny, nz, fc, ghost);
      double fp = kp >= 0 ? src[(int64_t)i * n + kp] : -1.0;
      double fm = km >= 0 ? src[(int64_t)i * n + km] : -1.0;
      double r;
      if (fp > 0.0 && fm > 0.0) r = (fp + fm) / 2.0;
      else if (fp > 0.0 && fm < 0.0) r = (fp + fabs(v)) / 2;
      else if (fp < 0.0 && fm > 0.0) r = (fm + fabs(v)) / 2;
      else r = fabs(v);
      fo[(int64_t)i * n + k] = r;
      nrep++;
    }
  }
  if (jacobi) free(src);
  return nrep;
}

int64_t o39_damp_periodic(int nx, int ny, int nz, double *fo, int shifted, int jacobi) {
  const int32_t fc[6] = {O39_FACE_PERIODIC, O39_FACE_PERIODIC, O39_FACE_PERIODIC,
                         O39_FACE_PERIODIC, O39_FACE_PERIODIC, O39_FACE_PERIODIC};
  return o39_damp_faces(nx, ny, nz, fc, fo, shifted, jacobi);
}

/* Source of population k pulled into node (a, b, cc): streaming.f95:365-474
 * in pull form.  A population whose upstream node x - c_k lies across a
 * non-periodic face is the one the reference's edge link fills (the link i
 * with x + c_i outside and k = opp(i) on the plain lattice, the link i = k
 * with x - c_i outside on the shifted one, where opp is the identity).  The
 * first face crossed in the chain order -x, -y, -z, +x, +y, +z decides:
 *   -1000  fi = fs_fieq(k), gi = fs_gieq(k)     freestream_edge_bc.f95:18-19
 *   -6     fi = fout(x, k)                      pure_outflow_bc.f95:17-20
 *   -10    fi = fout(x, opp(k))                 bounceback.f95:13-19
 *   -11    specular: the population leaves x' = x - c_k reflected about the
 *          half-way plane of every crossed inviscid face, with the mirrored
 *          velocity (inviscid_edge.f95:28-1243 tabulates the same remap as
 *          direction inviscid_dir written at x + loc; a face crossed after the
 *          reflection that is not inviscid then decides as above).
 *   -100-j / -200-j  inlet j / outlet j (inlet_bc.f95:21-53, outlet_bc.f95:
 *          23-58): face_row[f] >= 0 the row of the inlet / outlet table (the
 *          equilibrium of the stored state, in direction k), -2 the freestream
 *          pair (type 3, freestream_bc.f95), -1 pure outflow (type 6, as -6).
 * Returns 1 for the freestream value, 2 + row for a table row, else 0 with
 * (*src, *dir) to read. */
static int source39(int64_t a, int64_t b, int64_t cc, int k, int nx, int ny, int nz,
                    const int32_t *fc, const int32_t *face_row, const int *cx, const int *cy,
                    const int *cz, int shifted, int64_t *src, int *dir) {
  const int64_t n[3] = {nx, ny, nz};
  int64_t s[3] = {a - cx[k], b - cy[k], cc - cz[k]};
  int first = -1;
  for (int f = 0; f < 6 && first < 0; f++)
    if (axis_out(s[f % 3], n[f % 3], f) && fc[f] != O39_FACE_PERIODIC) first = f;
  for (int d = 0; d < 3; d++) {
    if (s[d] < 0 && fc[d] == O39_FACE_PERIODIC) s[d] += n[d];
    if (s[d] >= n[d] && fc[d + 3] == O39_FACE_PERIODIC) s[d] -= n[d];
  }
  int m[3] = {cx[k], cy[k], cz[k]};
  if (first >= 0 && fc[first] == -11) {
    for (int d = 0; d < 3; d++) {
      if (s[d] < 0 && fc[d] == -11) { s[d] = -1 - s[d]; m[d] = -m[d]; }
      else if (s[d] >= n[d] && fc[d + 3] == -11) { s[d] = 2 * n[d] - 1 - s[d]; m[d] = -m[d]; }
    }
    first = -1;
    for (int f = 0; f < 6 && first < 0; f++)
      if (axis_out(s[f % 3], n[f % 3], f)) first = f;
  }
  const int64_t own = a + n[0] * (b + n[1] * cc);
  if (first < 0) {
    int j = 0;
    while (j < NQ && !(cx[j] == m[0] && cy[j] == m[1] && cz[j] == m[2])) j++;
    *src = s[0] + n[0] * (s[1] + n[1] * s[2]);
    *dir = j;
    return 0;
  }
  if (fc[first] == -1000) return 1;
  if (fc[first] <= -101 && fc[first] >= -299) {
    const int r = face_row ? face_row[first] : -1;
    if (r == -2) return 1;
    if (r >= 0) return 2 + r;
  }
  *src = own;
  *dir = (fc[first] == -10 && !shifted) ? O39_OPP[k] : k;
  return 0;
}

void o39_init_fields(int nx, int ny, int nz, const double *rho, const double *ux, const double *uy,
                     const double *uz, const double *T, const o39_consts *c, double *f, double *g,
                     double *lm, int64_t *bad) {
  const int64_t n = (int64_t)nx * ny * nz;
  int64_t nb = 0;
#pragma omp parallel for schedule(static) reduction(+ : nb)
  for (int64_t k = 0; k < n; k++) {
    double u[3] = {ux[k], uy[k], uz[k]}, P[6], l[NM], e[NQ];
    P[0] = rho[k] * (T[k] + u[0] * u[0]);
    P[1] = rho[k] * (T[k] + u[1] * u[1]);
    P[2] = rho[k] * (T[k] + u[2] * u[2]);
    P[3] = rho[k] * u[0] * u[1];
    P[4] = rho[k] * u[0] * u[2];
    P[5] = rho[k] * u[1] * u[2];
    /* start from the Maxwellian's multipliers for rho = 1 at (u, T) */
    memset(l, 0, sizeof(l));
    l[0] = -1.5 * log(2 * C_PI * T[k]) - (u[0] * u[0] + u[1] * u[1] + u[2] * u[2]) / (2 * T[k]);
    l[1] = u[0] / T[k]; l[2] = u[1] / T[k]; l[3] = u[2] / T[k];
    l[4] = l[5] = l[6] = -1.0 / (2 * T[k]);
    int it = o39_lm_solve(rho[k], u, T[k], P, l, e, c->shifted, 0, 200);
    if (it < 0) nb++;
    o39_feq(l, rho[k], c->shifted, e);
    for (int i = 0; i < NQ; i++) {
      f[(int64_t)i * n + k] = e[i];
      g[(int64_t)i * n + k] = c->lead_g * T[k] * e[i];
    }
    for (int j = 0; j < NM; j++) lm[(int64_t)j * n + k] = l[j];
  }
  *bad = nb;
}

/* source39 on the lattice of `shifted`, for tests */
int o39_source(int64_t a, int64_t b, int64_t cc, int k, int nx, int ny, int nz, const int32_t fc[6],
               int shifted, int64_t *src, int *dir) {
  int cx[NQ], cy[NQ], cz[NQ];
  o39_velocities(shifted, cx, cy, cz);
  return source39(a, b, cc, k, nx, ny, nz, fc, NULL, cx, cy, cz, shifted, src, dir);
}

int o39_source_bc(int64_t a, int64_t b, int64_t cc, int k, int nx, int ny, int nz,
                  const int32_t fc[6], const int32_t *face_row, int shifted, int64_t *src,
                  int *dir) {
  int cx[NQ], cy[NQ], cz[NQ];
  o39_velocities(shifted, cx, cy, cz);
  return source39(a, b, cc, k, nx, ny, nz, fc, face_row, cx, cy, cz, shifted, src, dir);
}

/* One or more steps of a box whose faces carry the codes fc (see source39);
 * fs_f / fs_g are the freestream equilibria (freestream_basis_calcs.f90:66-71)
 * and may be NULL when no face is -1000. */
static void step_core(int nx, int ny, int nz, const int32_t fc[6], int ghost, const double *fs_f,
                      const double *fs_g, const int32_t *face_row, const double *bc_tab,
                      double *f, double *g, double *lm, double *alpha_out, const o39_consts *c,
                      int64_t step0, int nsteps, int ref_lu, int jacobi, o39_stats *st);

void o39_step_faces(int nx, int ny, int nz, const int32_t fc[6], const double *fs_f,
                    const double *fs_g, const int32_t *face_row, const double *bc_tab, double *f,
                    double *g, double *lm, double *alpha_out, const o39_consts *c, int64_t step0,
                    int nsteps, int ref_lu, int jacobi, o39_stats *st) {
  step_core(nx, ny, nz, fc, 0, fs_f, fs_g, face_row, bc_tab, f, g, lm, alpha_out, c, step0, nsteps,
            ref_lu, jacobi, st);
}

/* A refined level stored with a ghost layer of depth ghost: only the valid
 * nodes inside it are streamed and collided (their links stay inside the
 * stored box), the layer keeps its values (it is refilled from the coarser
 * level before the next step). */
void o39_step_ghost(int nx, int ny, int nz, int ghost, double *f, double *g, double *lm,
                    double *alpha_out, const o39_consts *c, int64_t step0, int nsteps, int ref_lu,
                    int jacobi, o39_stats *st) {
  const int32_t fc[6] = {-10, -10, -10, -10, -10, -10};
  step_core(nx, ny, nz, fc, ghost, NULL, NULL, NULL, NULL, f, g, lm, alpha_out, c, step0, nsteps,
            ref_lu, jacobi, st);
}

static void step_core(int nx, int ny, int nz, const int32_t fc[6], int ghost, const double *fs_f,
                      const double *fs_g, const int32_t *face_row, const double *bc_tab,
                      double *f, double *g, double *lm, double *alpha_out, const o39_consts *c,
                      int64_t step0, int nsteps, int ref_lu, int jacobi, o39_stats *st) {
  int cx[NQ], cy[NQ], cz[NQ];
  o39_velocities(c->shifted, cx, cy, cz);
  const int64_t n = (int64_t)nx * ny * nz;
  double *fo = (double *)malloc(sizeof(double) * (size_t)(n * NQ));
  double *go = (double *)malloc(sizeof(double) * (size_t)(n * NQ));
  memset(st, 0, sizeof(*st));
  for (int s = 0; s < nsteps; s++) {
    int64_t step = step0 + s;
    int64_t it_sum = 0, noconv = 0, a2 = 0, alo = 0, ahi = 0, nan = 0, kerr = 0;
    /* post-collision state of the previous step lives in f, g: stream it */
#pragma omp parallel for schedule(static) reduction(+ : it_sum, noconv, a2, alo, ahi, nan, kerr)
    for (int64_t k = 0; k < n; k++) {
      int64_t a = k % nx, b = (k / nx) % ny, cc = k / ((int64_t)nx * ny);
      double fi[NQ], gi[NQ], l[NM], e[NQ], u[3], P[6], rho, T;
      if (in_ghost(a, b, cc, nx, ny, nz, ghost)) {
        for (int i = 0; i < NQ; i++) {
          fo[(int64_t)i * n + k] = f[(int64_t)i * n + k];
          go[(int64_t)i * n + k] = g[(int64_t)i * n + k];
        }
        continue;
      }
      for (int i = 0; i < NQ; i++) {
        int64_t src;
        int dir;
        const int r = source39(a, b, cc, i, nx, ny, nz, fc, face_row, cx, cy, cz, c->shifted,
                               &src, &dir);
        if (r == 1) {
          fi[i] = fs_f[i];
          gi[i] = fs_g[i];
        } else if (r >= 2) {
          fi[i] = bc_tab[(int64_t)(r - 2) * 2 * NQ + i];
          gi[i] = bc_tab[(int64_t)(r - 2) * 2 * NQ + NQ + i];
        } else {
          fi[i] = f[(int64_t)dir * n + src];
          gi[i] = g[(int64_t)dir * n + src];
        }
      }
      o39_moments(fi, gi, c->shifted, &rho, u, &T, P, c->cv);
      for (int j = 0; j < NM; j++) l[j] = lm[(int64_t)j * n + k];
      int it = o39_lm_solve(rho, u, T, P, l, e, c->shifted, ref_lu, c->maxit);
      if (it < 0) { noconv++; it = c->maxit; }
      it_sum += it;
      for (int j = 0; j < NM; j++) lm[(int64_t)j * n + k] = l[j];
      double diff[NQ];
      for (int i = 0; i < NQ; i++) diff[i] = e[i] - fi[i];
      double alpha = 2.0;
      if (c->turbulent) {
        int stt;
        alpha = o39_alpha_estimate(fi, diff, T, step, c, &stt);
        if (stt == 1) a2++;
        else if (stt == 2) alo++;
        else if (stt == 3) ahi++;
        else if (stt < 0) nan++;
      }
      if (alpha_out) alpha_out[k] = alpha;
      double fq[NQ], gq[NQ];
      if (o39_collide(fi, gi, l, rho, u, T, alpha, step, c, fq, gq)) kerr++;
      for (int i = 0; i < NQ; i++) {
        fo[(int64_t)i * n + k] = fq[i];
        go[(int64_t)i * n + k] = gq[i];
      }
    }
    {
      int64_t valid = n;
      if (ghost > 0)
        valid = (int64_t)(nx - 2 * ghost) * (ny - 2 * ghost) * (nz - 2 * ghost);
      st->nodes += valid;
    }
    st->lm_iters += it_sum;
    st->lm_noconv += noconv;
    st->n_alpha2 += a2;
    st->n_alpha_lo += alo;
    st->n_alpha_hi += ahi;
    st->n_nan += nan;
    st->n_kappa_err += kerr;
    st->n_repair += damp_core(nx, ny, nz, fc, ghost, fo, c->shifted, jacobi);
    memcpy(f, fo, sizeof(double) * (size_t)(n * NQ));
    memcpy(g, go, sizeof(double) * (size_t)(n * NQ));
  }
  free(fo);
  free(go);
}

void o39_step_periodic(int nx, int ny, int nz, double *f, double *g, double *lm, double *alpha_out,
                       const o39_consts *c, int64_t step0, int nsteps, int ref_lu, int jacobi,
                       o39_stats *st) {
  const int32_t fc[6] = {O39_FACE_PERIODIC, O39_FACE_PERIODIC, O39_FACE_PERIODIC,
                         O39_FACE_PERIODIC, O39_FACE_PERIODIC, O39_FACE_PERIODIC};
  o39_step_faces(nx, ny, nz, fc, NULL, NULL, NULL, NULL, f, g, lm, alpha_out, c, step0, nsteps,
                 ref_lu, jacobi, st);
}

/* nrep LU updates on the same inputs (timing against the reference's
   lu_decomp): the sum of each call's max|delta| and first updated multiplier */
double o39_lu_batch(int nrep, const double *J, const double *mom, const double *lm0, int ref_mode) {
  double s = 0.0, lm[13];
  for (int k = 0; k < nrep; ++k) {
    memcpy(lm, lm0, sizeof lm);
    const double mx = o39_lu_step(J, mom, lm, ref_mode);
    s = (s + mx) + lm[0];
  }
  return s;
}

/* o39_sutherland + o39_therm_cond at n nodes (timing): the sum of mu + kappa */
double o39_quick_batch(int n, const double *rho, const double *T, const o39_consts *c) {
  double s = 0.0, dm, dt, kappa;
  for (int k = 0; k < n; ++k) {
    s += o39_sutherland(T[k], c, &dm, &dt);
    if (o39_therm_cond(rho[k], T[k], c, &kappa) == 0) s += kappa;
  }
  return s;
}
