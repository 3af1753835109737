/* D3Q39 step of a box with a shifted-lattice link table: the reference's
 * curved-wall rule for the shifted lattice, FP64.  TEST INFRASTRUCTURE ONLY
 * (see uclbs_oracle.h).
 *
 * A fluid node x whose link i (keyed on the upstream node x - c_i, see
 * o_build_links39s) carries a wall code s <= -1001 has its population i
 * rebuilt instead of pulled from the solid upstream node
 * (looping/timestepping/streaming.f95:478-1133 with the shifted branch, where
 * opp is the identity, so the rebuilt value is f_i(x) itself):
 *
 *   fields:  rho, u, T of a node are the moments of its populations in the
 *     input buffers (post-collision of the previous step; the reference reads
 *     the fields of the previous step's calculate_flow_variables, which the
 *     collision conserves up to the q-bar term on the shifted lattice).
 *   target (target_finder.f95:87-127), when a link of speed <= sqrt(3)
 *     (Fortran directions 2..15) is a wall link (streaming.f95:491, 749):
 *       u_t = mean over wall links i of q_i u_n / (q_i + 1),
 *       T_t = mean of (q_i T_n + T_n) / (q_i + 1),
 *     n the downstream node x + c_i when the real opposite link erp(i) is
 *     null (-666), else x itself (the reference's u_pass(5): Fortran
 *     direction 5 is the shifted lattice's (0,0,0));
 *       rho_t = sum over directions of f_erp(i)(x) where link i is a wall
 *     link, else f_i(x) (this step's pulled values: the reference sums the fi
 *     array while streaming it direction by direction, a mix of this and the
 *     previous step's values; the order-free form is used here);
 *     target multipliers by the multiplier Newton on (rho_t, u_t, T_t,
 *     P = rho_t (T_t d + u_t u_t), q = (rho_t |u_t|^2 + 5 rho_t T_t) u_t)
 *     from the node's multipliers, or from the fixed start-up guesses of
 *     freestream_basis_calcs.f90:154-189 while the level's own step counter
 *     is <= 4 (streaming.f95:777-795).
 *   otherwise (only longer links hit the wall): the node's own rho, u, T and
 *     stored multipliers (streaming.f95:1051-1084; for g the node's T -- the
 *     reference's shifted branch reads a target temperature never set for
 *     such a node).
 *   derivatives (derivatives.f95:816-1236, 1236-...): along axis a of wall
 *     link i, the stencil kind of the reference's case table (+: (0 - u_t) /
 *     (1 + w), -: (u_t - 0) / (1 + w), 0: the local difference of
 *     u(x - e_a), u(x), u(x + e_a) over the fluid ones): the sign of the
 *     link's unshifted component, the shifted rest direction taking the +
 *     kind along x, and the reversed kinds for the four links (2,0,-2),
 *     (-2,0,2), (0,2,-2), (0,-2,2);
 *     w = |s_a + 1001| 1e-9 |c_i| with s_a the node code of x + e_a (the
 *     reference's down_state); temperature stencils of the wall kinds are 0.
 *   rebuilt populations (shifted_ibb.f95:48-269):
 *     f_i = feq_i(target) + eps k_f (P_comp + Q_comp), k_f = -rho T_t / 2,
 *     g_i = (2 c_v - 3) T_t feq_i(target) + eps k_g (3 k1 + g_a + R_comp),
 *     with the reference's 27-entry P, Q, R tensors of the velocity and
 *     temperature gradients, beta_1, beta_2 and eps (start-up ramps on the
 *     finest level's counter), including its term slips: P_zzx counted
 *     twice in P_comp, and the zy entries of R repeating the yy ones; a
 *     negative result is replaced by its magnitude (:262-267).
 * A node outside the box (a stencil neighbour beyond a non-periodic face)
 * counts as not fluid.  The shifted rest direction has no link code: it is
 * pulled from x - (1,0,0) whatever that node is (a solid node holds its
 * initial state).  Solid nodes keep their values.  The positivity repair
 * follows damp_it.f95:125-177 with the table: a neighbour along +c_i counts
 * when the link code is >= -500, along -c_i when that node's code is >= -500;
 * with neither, the population is left as it is. */
#include "uclbs_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define NQ 39
#define NM 13
#define O39_FACE_PERIODIC 2147483647

typedef struct { double rho, u[3], T; } macro39;

static macro39 node_macro(const double *f, const double *g, int64_t n, int64_t k,
                          const o39_consts *c) {
  double fl[NQ], gl[NQ], P[6];
  for (int i = 0; i < NQ; i++) {
    fl[i] = f[(int64_t)i * n + k];
    gl[i] = g[(int64_t)i * n + k];
  }
  macro39 m;
  o39_moments(fl, gl, c->shifted, &m.rho, m.u, &m.T, P, c->cv);
  return m;
}

/* node index of position (a, b, cc), wrapped over periodic faces; -1 beyond
 * another face */
static int64_t pos_index(int64_t a, int64_t b, int64_t cc, int nx, int ny, int nz,
                         const int32_t fc[6]) {
  int64_t s[3] = {a, b, cc};
  const int64_t nd[3] = {nx, ny, nz};
  for (int d = 0; d < 3; d++) {
    if (s[d] < 0) {
      if (fc[d] != O39_FACE_PERIODIC) return -1;
      s[d] += nd[d];
    } else if (s[d] >= nd[d]) {
      if (fc[d + 3] != O39_FACE_PERIODIC) return -1;
      s[d] -= nd[d];
    }
  }
  return s[0] + nd[0] * (s[1] + nd[1] * s[2]);
}

/* stencil kind of direction i along axis a (derivatives.f95:816-1235, the
 * case table of first_order_shifted_vel_cond_unknown; the temperature
 * function at :1236 has the same kinds): 1 where the case calls dudx_ibb with
 * winds = .false., -1 with winds = .true., 0 where it calls dudx_local.  It
 * is the sign of the link's unshifted component, the rest direction taking
 * the + kind along x, except for (2,0,-2), (-2,0,2), (0,2,-2) and (0,-2,2)
 * (Fortran directions 28, 30, 31, 33), whose cases carry the reversed kinds. */
static const signed char WALL_KIND[NQ][3] = {
    {1, 0, 0},   {1, 0, 0},  {0, 1, 0},  {0, 0, 1},   {-1, 0, 0},  {0, -1, 0}, {0, 0, -1},
    {1, 1, 1},   {-1, 1, 1}, {-1, -1, 1}, {1, -1, 1}, {1, 1, -1},  {-1, 1, -1}, {-1, -1, -1},
    {1, -1, -1}, {1, 0, 0},  {0, 1, 0},  {0, 0, 1},   {-1, 0, 0},  {0, -1, 0}, {0, 0, -1},
    {1, 1, 0},   {1, 0, 1},  {0, 1, 1},  {-1, 1, 0},  {-1, -1, 0}, {1, -1, 0}, {-1, 0, 1},
    {-1, 0, -1}, {1, 0, -1}, {0, -1, 1}, {0, -1, -1}, {0, 1, -1},  {1, 0, 0},  {0, 1, 0},
    {0, 0, 1},   {-1, 0, 0}, {0, -1, 0}, {0, 0, -1}};

static int dkind(int i, int a) { return WALL_KIND[i][a]; }

/* dudx_local (derivatives.f95:405-418) */
static double d_local(double rest, double up, double down, int up_ok, int down_ok) {
  if (up_ok && down_ok) return (down - up) / 2.0;
  if (up_ok) return rest - up;
  if (down_ok) return down - rest;
  return 0.0;
}

/* beta_1_calculator / beta_2_calculator (amr_processes.f95:1751-1785) */
static void betas(double mu, double kappa, double rho, double T, int64_t step,
                  const o39_consts *c, double *b1, double *b2) {
  if (step < c->startup) {
    *b1 = c->init_beta1 + c->beta1_scale * (double)step;
    *b2 = c->init_beta2 + c->beta2_scale * (double)step;
  } else {
    *b1 = 1.0 / ((ldexp(1.0, c->lvl + 1) * mu) / (rho * T) + 1.0);
    *b2 = 1.0 / ((2.0 * kappa) / (rho * c->cp * T) + 1.0);
  }
}

/* start-up multiplier guesses (freestream_basis_calcs.f90:154-189) */
static void startup_lm(double mach, double lm[NM]) {
  static const double hi[NM] = {-2.2, 0.4, 0.0, 0.0, -0.45, -0.45, -0.45, 0.0, 0.0, 0.0,
                                0.015, 0.0, 0.0};
  static const double lo[NM] = {-2.0, 0.0, 0.0, 0.0, -0.4, -0.4, -0.4, 0.0, 0.0, 0.0,
                                0.002, 0.0, 0.0};
  memcpy(lm, mach > 1.0 ? hi : lo, sizeof(double) * NM);
}

/* Velocity and temperature gradients at a wall node for wall link i
 * (derivatives.f95:816-1236): su/sT/scode hold u, T and the node code of the
 * 7-point stencil x, x + e_x, x + e_y, x + e_z, x - e_x, x - e_y, x - e_z (a
 * code < 0 marks a stencil node that is not fluid).  Along axis a the kind is
 * WALL_KIND[i][a]; + kind: (0 - u_t) / (1 + w), - kind: (u_t - 0) / (1 + w),
 * with w = |s_a + 1001| 1e-9 |c_i| from the code s_a of x + e_a; kind 0: the
 * central difference over fluid neighbours, else the one-sided one, else 0.
 * Temperature gradients of the wall kinds are 0.  du[a][b] = d u_b / d x_a. */
void o39_wall_derivs(int i, const double ut[3], double su[7][3], const double sT[7],
                     const int scode[7], int shifted, double du[3][3], double dT[3]) {
  int cx[NQ], cy[NQ], cz[NQ];
  o39_velocities(shifted, cx, cy, cz);
  const double link = sqrt((double)(cx[i] * cx[i] + cy[i] * cy[i] + cz[i] * cz[i]));
  for (int ax = 0; ax < 3; ax++) {
    const int kind = dkind(i, ax);
    const double wd = fabs((double)scode[1 + ax] + 1001.0) / 1.0e9 * link;
    const int up_ok = scode[4 + ax] >= 0, down_ok = scode[1 + ax] >= 0;
    for (int bb = 0; bb < 3; bb++) {
      if (kind > 0) du[ax][bb] = (0.0 - ut[bb]) / (1.0 + wd);
      else if (kind < 0) du[ax][bb] = (ut[bb] - 0.0) / (1.0 + wd);
      else du[ax][bb] = d_local(su[0][bb], su[4 + ax][bb], su[1 + ax][bb], up_ok, down_ok);
    }
    dT[ax] = kind != 0 ? 0.0 : d_local(sT[0], sT[4 + ax], sT[1 + ax], up_ok, down_ok);
  }
}

/* The rebuilt f and g of wall link i (shifted_ibb.f95:48-259) before the
 * magnitude repair: target rho_t, u_t, T_t with feq_i the target equilibrium
 * of direction i, gradients du[a][b] = d u_b / d x_a and dT, ramped eps and
 * beta_1, beta_2. */
void o39_wall_grad(int i, double rho_t, const double ut[3], double Tt, double feq_i,
                   double du[3][3], const double dT[3], double eps, double b1, double b2,
                   const o39_consts *c, double *f_out, double *g_out) {
  int cx[NQ], cy[NQ], cz[NQ];
  o39_velocities(c->shifted, cx, cy, cz);
  const double cv = c->cv, T = Tt, u = ut[0], v = ut[1], w = ut[2];
  const double k_f = -(rho_t * T) / 2.0, k_g = -(rho_t * T * (2.0 * cv - 3.0)) / 2.0;
  const double dudx = du[0][0], dudy = du[1][0], dudz = du[2][0];
  const double dvdx = du[0][1], dvdy = du[1][1], dvdz = du[2][1];
  const double dwdx = du[0][2], dwdy = du[1][2], dwdz = du[2][2];
  const double dTdx = dT[0], dTdy = dT[1], dTdz = dT[2];
  double p[28], qq[28], r[28]; /* 1-based as in shifted_ibb.f95 */
  p[1] = 2 * dudx - dudx / cv; p[2] = 2 * dudx - dvdy / cv; p[3] = 2 * dudx - dwdz / cv;
  p[4] = p[5] = p[6] = dudy + dvdx;
  p[7] = p[8] = p[9] = dudz + dwdx;
  p[10] = p[11] = p[12] = dudy + dvdx;
  p[13] = 2 * dvdy - dudx / cv; p[14] = 2 * dvdy - dvdy / cv; p[15] = 2 * dvdy - dwdz / cv;
  p[16] = p[17] = p[18] = dvdz + dwdy;
  p[19] = p[20] = p[21] = dudz + dwdx;
  p[22] = p[23] = p[24] = dvdz + dwdy;
  p[25] = 2 * dwdz - dudx / cv; p[26] = 2 * dwdz - dvdy / cv; p[27] = 2 * dwdz - dwdz / cv;
  for (int j = 1; j <= 27; j++) p[j] = p[j] / b1;
  const double tx = dTdx / b2, ty = dTdy / b2, tz = dTdz / b2;
  qq[1] = 3 * (tx + u * p[1]);
  qq[2] = ty + v * p[2] + 2 * u * p[4];
  qq[3] = tz + w * p[3] + 2 * u * p[7];
  qq[4] = ty + 2 * u * p[4] + v * p[1];
  qq[5] = tx + u * p[13] + 2 * v * p[4];
  qq[6] = u * p[22] + v * p[7] + w * p[4];
  qq[7] = tz + w * p[3] + 2 * u * p[7];
  qq[8] = u * p[22] + v * p[7] + w * p[4];
  qq[9] = tx + u * p[25] + 2 * w * p[7];
  qq[10] = ty + v * p[2] + 2 * u * p[4];
  qq[11] = tx + 2 * v * p[4] + u * p[13];
  qq[12] = u * p[22] + v * p[7] + w * p[4];
  qq[13] = tx + u * p[13] + 2 * v * p[4];
  qq[14] = 3 * (ty + v * p[14]);
  qq[15] = tz + w * p[15] + 2 * v * p[22];
  qq[16] = u * p[22] + v * p[7] + w * p[4];
  qq[17] = tz + w * p[15] + 2 * v * p[22];
  qq[18] = ty + v * p[26] + 2 * w * p[22];
  qq[19] = tx + w * p[3] + 2 * u * p[7];
  qq[20] = u * p[22] + v * p[7] + w * p[4];
  qq[21] = tx + u * p[25] + 2 * w * p[7];
  qq[22] = u * p[22] + v * p[7] + w * p[4];
  qq[23] = tz + w * p[15] + 2 * v * p[22];
  qq[24] = ty + v * p[26] + 2 * w * p[22];
  qq[25] = tx + u * p[25] + 2 * w * p[7];
  qq[26] = ty + v * p[26] + 2 * w * p[22];
  qq[27] = 3 * (tz + w * p[27]);
  const double X = cx[i], Y = cy[i], Z = cz[i];
  /* second-order basis of the 27 (a, b, *) entries */
  const double sxx = X * X - T, syy = Y * Y - T, szz = Z * Z - T;
  const double pc[28] = {0,   sxx, sxx, sxx, X * Y, X * Y, X * Y, X * Z, X * Z, X * Z,
                         Y * X, Y * X, Y * X, syy, syy, syy, Y * Z, Y * Z, Y * Z,
                         Z * X, Z * X, Z * X, Z * Y, Z * Y, Z * Y, szz, szz, szz};
  double psum = 0.0;
  for (int j = 1; j <= 27; j++) psum += p[j] * pc[j];
  psum += p[25] * szz; /* P_zzx counted twice (shifted_ibb.f95:170) */
  const double p_comp = psum / (2 * T * T);
  const double qc[28] = {0,
                         X * X * X - 3 * X * T, X * X * Y - 3 * Y * T, X * X * Z - 3 * Z * T,
                         X * X * Y, X * Y * Y, X * Y * Z, X * X * Z, X * Y * Z, X * Z * Z,
                         X * X * Y, Y * Y * X, X * Y * Z,
                         Y * Y * X - 3 * X * T, Y * Y * Y - 3 * Y * T, Y * Y * Z - 3 * Z * T,
                         X * Y * Z, Y * Y * Z, Z * Z * Y, X * X * Z, X * Y * Z, Z * Z * X,
                         X * Y * Z, Z * Y * Y, Z * Z * Y,
                         Z * Z * X - 3 * X * T, Z * Z * Y - 3 * Y * T, Z * Z * Z - 3 * Z * T};
  double qsum = 0.0;
  for (int j = 1; j <= 27; j++) qsum += qq[j] * qc[j];
  const double q_comp = qsum / (6 * T * T * T);
  const double fo = feq_i + eps * k_f * (p_comp + q_comp);
  const double k_one = 3 * (dTdx + dTdy + dTdz) / (b1 * cv);
  const double g_alpha = X / T * (9 * dTdx / b2 + u * k_one) + Y / T * (9 * dTdy / b2 + v * k_one) +
                         Z / T * (9 * dTdz / b2 + w * k_one);
  r[1] = 2 * T * dudx + 2 * u * dTdx - dudx / cv * (2 * T + u * u);
  r[2] = 2 * T * dudx + 2 * u * dTdx - dvdy / cv * (2 * T + u * u);
  r[3] = 2 * T * dudx + 2 * u * dTdx - dwdz / cv * (2 * T + u * u);
  r[4] = T * (dudy + dvdx) + u * dTdy + v * dTdx - dudx / cv * u * v;
  r[5] = T * (dudy + dvdx) + u * dTdy + v * dTdx - dvdy / cv * u * v;
  r[6] = T * (dudy + dvdx) + u * dTdy + v * dTdx - dwdz / cv * u * v;
  r[7] = T * (dudz + dwdx) + u * dTdz + w * dTdx - dudx / cv * u * w;
  r[8] = T * (dudz + dwdx) + u * dTdz + w * dTdx - dvdy / cv * u * w;
  r[9] = T * (dudz + dwdx) + u * dTdz + w * dTdx - dwdz / cv * u * w;
  r[10] = r[4]; r[11] = r[5]; r[12] = r[6];
  r[13] = 2 * T * dvdy + 2 * v * dTdy - dudx / cv * (2 * T + v * v);
  r[14] = 2 * T * dvdy + 2 * v * dTdy - dvdy / cv * (2 * T + v * v);
  r[15] = 2 * T * dvdy + 2 * v * dTdy - dwdz / cv * (2 * T + v * v);
  r[16] = T * (dvdz + dwdy) + w * dTdy + v * dTdz - dudx / cv * v * w;
  r[17] = T * (dvdz + dwdy) + w * dTdy + v * dTdz - dvdy / cv * v * w;
  r[18] = T * (dvdz + dwdy) + w * dTdy + v * dTdz - dwdz / cv * v * w;
  r[19] = r[7]; r[20] = r[8]; r[21] = r[9];
  r[22] = r[13]; r[23] = r[14]; r[24] = r[15]; /* the zy entries repeat yy (:231-233) */
  r[25] = 2 * T * dwdz + 2 * w * dTdz - dudx / cv * (2 * T + w * w);
  r[26] = 2 * T * dwdz + 2 * w * dTdz - dvdy / cv * (2 * T + w * w);
  r[27] = 2 * T * dwdz + 2 * w * dTdz - dwdz / cv * (2 * T + w * w);
  double rsum = 0.0;
  for (int j = 1; j <= 27; j++) rsum += r[j] * pc[j];
  const double gi_neq = k_g * (3 * k_one + g_alpha + rsum / (2 * b1 * T));
  const double go = c->lead_g * T * feq_i + eps * gi_neq;
  *f_out = fo;
  *g_out = go;
}

/* The rebuilt f and g of the wall links of fluid node k (positions a, b, cc).
 * fi: this step's pulled populations of the node (wall links: the value
 * pulled from the solid upstream node).  Writes fw/gw for the wall links;
 * returns the number of magnitude repairs; *noconv / *kerr flag a target
 * Newton at its cap and a conductivity outside its fit. */
int o39_wall_rebuild(int nx, int ny, int nz, const int32_t fc[6], const int32_t *state,
                     const double *f, const double *g, const double *lm, int64_t k,
                     const double fi[NQ], const o39_consts *c, int64_t step, int64_t lvl_step,
                     double fw[NQ], double gw[NQ], int *noconv, int *kerr) {
  const int64_t n = (int64_t)nx * ny * nz;
  int cx[NQ], cy[NQ], cz[NQ];
  o39_velocities(c->shifted, cx, cy, cz);
  const int64_t a = k % nx, b = (k / nx) % ny, cc = k / ((int64_t)nx * ny);
  const macro39 own = node_macro(f, g, n, k, c);
  int adjacent = 0;
  for (int j = 1; j <= 14; j++) adjacent |= state[(int64_t)j * n + k] < -1000;
  double rho_t, ut[3], Tt, lmt[NM], feq[NQ];
  *noconv = 0;
  *kerr = 0;
  if (adjacent) {
    int count = 0;
    ut[0] = ut[1] = ut[2] = 0.0;
    Tt = 0.0;
    for (int i = 1; i < NQ; i++) {
      const int32_t s = state[(int64_t)i * n + k];
      if (s > -1001) continue;
      const double q = fabs((double)s + 1001.0) / 1.0e9;
      count++;
      macro39 nb = own;
      if (state[(int64_t)O39_ERP[i] * n + k] == -666) {
        const int64_t d = pos_index(a + cx[i], b + cy[i], cc + cz[i], nx, ny, nz, fc);
        if (d >= 0) nb = node_macro(f, g, n, d, c);
      }
      for (int e = 0; e < 3; e++) ut[e] += (q * nb.u[e]) / (q + 1.0);
      Tt += (q * nb.T + nb.T) / (q + 1.0);
    }
    for (int e = 0; e < 3; e++) ut[e] /= count;
    Tt /= count;
    rho_t = 0.0;
    for (int i = 0; i < NQ; i++)
      rho_t += state[(int64_t)i * n + k] < -1000 ? fi[O39_ERP[i]] : fi[i];
    if (lvl_step <= 4) startup_lm(c->mach, lmt);
    else
      for (int j = 0; j < NM; j++) lmt[j] = lm[(int64_t)j * n + k];
    double P[6] = {rho_t * Tt + rho_t * ut[0] * ut[0], rho_t * Tt + rho_t * ut[1] * ut[1],
                   rho_t * Tt + rho_t * ut[2] * ut[2], rho_t * ut[0] * ut[1],
                   rho_t * ut[0] * ut[2], rho_t * ut[1] * ut[2]};
    double e[NQ];
    if (o39_lm_solve(rho_t, ut, Tt, P, lmt, e, c->shifted, 0, c->maxit) < 0) *noconv = 1;
  } else {
    rho_t = own.rho;
    for (int e = 0; e < 3; e++) ut[e] = own.u[e];
    Tt = own.T;
    for (int j = 0; j < NM; j++) lmt[j] = lm[(int64_t)j * n + k];
  }
  o39_feq(lmt, rho_t, c->shifted, feq);
  /* 7-point stencil: x, x + e_a, x - e_a */
  static const int ex[7][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1},
                               {-1, 0, 0}, {0, -1, 0}, {0, 0, -1}};
  double su[7][3], sT[7];
  int scode[7];
  for (int j = 0; j < 7; j++) {
    const int64_t d = pos_index(a + ex[j][0], b + ex[j][1], cc + ex[j][2], nx, ny, nz, fc);
    scode[j] = d >= 0 ? state[d] : -1;
    if (scode[j] >= 0) {
      const macro39 m = j == 0 ? own : node_macro(f, g, n, d, c);
      for (int e = 0; e < 3; e++) su[j][e] = m.u[e];
      sT[j] = m.T;
    } else {
      su[j][0] = su[j][1] = su[j][2] = 0.0;
      sT[j] = 1.0 / c->gama;
    }
  }
  const double eps =
      step < c->startup ? (((double)step + 1.0) / (double)c->startup) * c->eps : c->eps;
  const double mu = o39_sutherland(Tt, c, NULL, NULL);
  double kappa = 0.0;
  if (o39_therm_cond(rho_t, Tt, c, &kappa)) *kerr = 1;
  double b1, b2;
  betas(mu, kappa, rho_t, Tt, step, c, &b1, &b2);
  int nrep = 0;
  for (int i = 1; i < NQ; i++) {
    if (state[(int64_t)i * n + k] > -1001) continue;
    double du[3][3], dT[3];
    o39_wall_derivs(i, ut, su, sT, scode, c->shifted, du, dT);
    double fo, go;
    o39_wall_grad(i, rho_t, ut, Tt, feq[i], du, dT, eps, b1, b2, c, &fo, &go);
    if (fo < 0.0) { fo = fabs(fo); nrep++; }
    if (go < 0.0) { go = fabs(go); nrep++; }
    fw[i] = fo;
    gw[i] = go;
  }
  return nrep;
}

/* ---- plain lattice (Mach < 1.2) ----------------------------------------------
 *
 * A fluid node x whose link j (x -> x + c_j) carries a wall code s <= -1001
 * has the population of the opposite direction i = opp(j), which the pull
 * would read from the solid node x + c_j, rebuilt
 * (streaming.f95:478-1133 with the plain branch, which calls
 * interpolative_bounceback_bc with direction i and stores the result in
 * fi(x, i)):
 *   target (target_finder.f95:36-88), when a link of speed <= sqrt(3) is a
 *     wall link: over the wall links j with q_j = |s_j + 1001| 1e-9,
 *       u_t = mean of q_j u_n / (q_j + 1), T_t = mean of (q_j T_n + T_n) /
 *       (q_j + 1), n the node x + c_opp(j) when the opposite link's code is
 *       >= 0 (the node's own values where that node lies beyond a face);
 *       otherwise u_t += q_j u(x) / (q_j + 1), T_t += (q_j T(x) + T_inf) /
 *       (q_j + 1), T_inf = 1 / gamma (the freestream temperature);
 *       rho_t = sum over directions of f_opp(i)(x) where link opp(i) is a
 *       wall link, else f_i(x) (this step's pulled values);
 *     target multipliers by Newton from the node's multipliers (the plain
 *     branch has no start-up guesses);
 *   otherwise the node's own rho, u, T and multipliers;
 *   derivatives (interpolative_bounceback_bc.f95:834-2838 with
 *     derivatives.f95:237-583): along axis a with c_i,a > 0 the wall form
 *     (u_t - 0) / (1 + w), c_i,a < 0: (0 - u_t) / (1 + w), with
 *     w = |s_i + 1001| 1e-9 |c_i| from the node's own code of link i (the
 *     direction of the rebuilt population, as the reference reads it);
 *     temperature derivatives of the wall form are 0; c_i,a = 0: the local
 *     difference of u(x - e_a), u(x), u(x + e_a) over those whose link code
 *     from x is >= 0 (the stencil values: a fluid node's own, 0 and T_inf
 *     for a node that is not fluid or lies beyond a face);
 *   rebuilt populations: f_i = feq_i(target) + eps k_f F_i,
 *     g_i = (2 c_v - 3) T_t feq_i + eps k_g G_i, k_f = -rho_t T_t / 2,
 *     k_g = -rho_t T_t (2 c_v - 3) / 2, where F_i, G_i are the Hermite
 *     projection the reference's 38 cases expand plus each case's slip
 *     terms (a table: o39p_set_slips); no magnitude repair (the plain branch
 *     has none; damp_it repairs negative populations after the collision).
 */
static const double *g_slip = NULL;
static const int *g_slip_start = NULL;

void o39p_set_slips(const double *slip, const int *start) {
  g_slip = slip;
  g_slip_start = start;
}

static double ipow(double x, int e) {
  double r = 1.0;
  for (int k = 0; k < (e < 0 ? -e : e); k++) r *= x;
  return e < 0 ? 1.0 / r : r;
}

/* F_i and G_i: the Hermite projection of direction i (1..38) plus its slips */
void o39p_wall_neq(int i, double T, const double ut[3], double du[3][3], const double dT[3],
                   double b1, double b2, double cv, double *F, double *G) {
  const int c[3] = {O39_CX[i], O39_CY[i], O39_CZ[i]};
  const double div = du[0][0] + du[1][1] + du[2][2];
  double h2p = 0.0, h2r = 0.0, h3q = 0.0;
  for (int a = 0; a < 3; a++)
    for (int b = 0; b < 3; b++) {
      const double h2 = (double)(c[a] * c[b]) - (a == b ? T : 0.0);
      double P, R;
      if (a == b) {
        P = 6.0 * du[a][a] - div / cv;
        R = 6.0 * T * du[a][a] + 6.0 * ut[a] * dT[a] - (2.0 * T + ut[a] * ut[a]) * div / cv;
      } else {
        P = 3.0 * (du[a][b] + du[b][a]);
        R = 3.0 * (T * (du[a][b] + du[b][a]) + ut[a] * dT[b] + ut[b] * dT[a]) -
            div / cv * ut[a] * ut[b];
      }
      h2p += h2 * P;
      h2r += h2 * R;
      for (int e = 0; e < 3; e++) {
        const double h3 = (double)(c[a] * c[b] * c[e]) -
                          T * (double)(c[a] * (b == e) + c[b] * (a == e) + c[e] * (a == b));
        if (h3 == 0.0) continue;
        /* Q_abe: the three ways to split {a, b, e} into a pair and a single */
        const int sp[3][3] = {{a, b, e}, {a, e, b}, {b, e, a}};
        double q = 0.0;
        for (int m = 0; m < 3; m++) {
          const int p = sp[m][0], r = sp[m][1], s = sp[m][2];
          const double pi = p == r ? 2.0 * du[p][p] - du[s][s] / cv : du[p][r] + du[r][p];
          q += (p == r ? dT[s] / b2 : 0.0) + ut[s] * pi / b1;
        }
        h3q += h3 * q;
      }
    }
  double f = h2p / (2.0 * b1 * T * T) + h3q / (6.0 * T * T * T);
  double gg = 9.0 * div / (b1 * cv) + h2r / (2.0 * b1 * T * T);
  for (int a = 0; a < 3; a++)
    gg += (double)c[a] * (9.0 * dT[a] / b2 + 3.0 * ut[a] * div / (b1 * cv)) / T;
  if (g_slip) {
    const double D[12] = {du[0][0], du[1][0], du[2][0], du[0][1], du[1][1], du[2][1],
                          du[0][2], du[1][2], du[2][2], dT[0], dT[1], dT[2]};
    for (int part = 0; part < 2; part++) {
      const int lo = g_slip_start[38 * part + i - 1], hi = g_slip_start[38 * part + i];
      double add = 0.0;
      for (int m = lo; m < hi; m++) {
        const double *t = g_slip + 8 * (size_t)m;
        add += t[0] * D[(int)t[1]] / ((int)t[2] == 1 ? b1 : b2) * ipow(ut[0], (int)t[3]) *
               ipow(ut[1], (int)t[4]) * ipow(ut[2], (int)t[5]) * ipow(T, (int)t[6]) *
               ipow(cv, (int)t[7]);
      }
      if (part == 0) f += add;
      else gg += add;
    }
  }
  *F = f;
  *G = gg;
}

/* The rebuilt f and g of the plain-lattice wall links of fluid node k:
 * fw[i], gw[i] for every direction i whose opposite link is a wall link.
 * fi: this step's pulled populations of the node. */
int o39p_wall_rebuild(int nx, int ny, int nz, const int32_t fc[6], const int32_t *state,
                      const double *f, const double *g, const double *lm, int64_t k,
                      const double fi[NQ], const o39_consts *c, int64_t step, double fw[NQ],
                      double gw[NQ], int *noconv, int *kerr) {
  const int64_t n = (int64_t)nx * ny * nz;
  const int64_t a = k % nx, b = (k / nx) % ny, cc = k / ((int64_t)nx * ny);
  const macro39 own = node_macro(f, g, n, k, c);
  const double t_inf = 1.0 / c->gama;
  int adjacent = 0;
  for (int j = 1; j <= 14; j++) adjacent |= state[(int64_t)j * n + k] < -1000;
  double rho_t, ut[3], Tt, lmt[NM], feq[NQ];
  *noconv = 0;
  *kerr = 0;
  if (adjacent) {
    int count = 0;
    ut[0] = ut[1] = ut[2] = 0.0;
    Tt = 0.0;
    for (int j = 1; j < NQ; j++) {
      const int32_t s = state[(int64_t)j * n + k];
      if (s > -1001) continue;
      const double q = fabs((double)s + 1001.0) / 1.0e9;
      count++;
      const int o = O39_OPP[j];
      if (state[(int64_t)o * n + k] >= 0) {
        macro39 nb = own;
        const int64_t d = pos_index(a + O39_CX[o], b + O39_CY[o], cc + O39_CZ[o], nx, ny, nz, fc);
        if (d >= 0) nb = node_macro(f, g, n, d, c);
        for (int e = 0; e < 3; e++) ut[e] += (q * nb.u[e]) / (q + 1.0);
        Tt += (q * nb.T + nb.T) / (q + 1.0);
      } else {
        for (int e = 0; e < 3; e++) ut[e] += (q * own.u[e]) / (q + 1.0);
        Tt += (q * own.T + t_inf) / (q + 1.0);
      }
    }
    for (int e = 0; e < 3; e++) ut[e] /= count;
    Tt /= count;
    rho_t = 0.0;
    for (int i = 0; i < NQ; i++)
      rho_t += state[(int64_t)O39_OPP[i] * n + k] < -1000 ? fi[O39_OPP[i]] : fi[i];
    for (int m = 0; m < NM; m++) lmt[m] = lm[(int64_t)m * n + k];
    double P[6] = {rho_t * Tt + rho_t * ut[0] * ut[0], rho_t * Tt + rho_t * ut[1] * ut[1],
                   rho_t * Tt + rho_t * ut[2] * ut[2], rho_t * ut[0] * ut[1],
                   rho_t * ut[0] * ut[2], rho_t * ut[1] * ut[2]};
    double e[NQ];
    if (o39_lm_solve(rho_t, ut, Tt, P, lmt, e, 0, 0, c->maxit) < 0) *noconv = 1;
  } else {
    rho_t = own.rho;
    for (int e = 0; e < 3; e++) ut[e] = own.u[e];
    Tt = own.T;
    for (int m = 0; m < NM; m++) lmt[m] = lm[(int64_t)m * n + k];
  }
  o39_feq(lmt, rho_t, 0, feq);
  /* 7-point stencil values x, x + e_a, x - e_a */
  static const int ex[7][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1},
                               {-1, 0, 0}, {0, -1, 0}, {0, 0, -1}};
  double su[7][3], sT[7];
  for (int m = 0; m < 7; m++) {
    const int64_t d = pos_index(a + ex[m][0], b + ex[m][1], cc + ex[m][2], nx, ny, nz, fc);
    if (d >= 0 && state[d] >= 0) {
      const macro39 mm = m == 0 ? own : node_macro(f, g, n, d, c);
      for (int e = 0; e < 3; e++) su[m][e] = mm.u[e];
      sT[m] = mm.T;
    } else {
      su[m][0] = su[m][1] = su[m][2] = 0.0;
      sT[m] = t_inf;
    }
  }
  const double eps =
      step < c->startup ? (((double)step + 1.0) / (double)c->startup) * c->eps : c->eps;
  const double mu = o39_sutherland(Tt, c, NULL, NULL);
  double kappa = 0.0;
  if (o39_therm_cond(rho_t, Tt, c, &kappa)) *kerr = 1;
  double b1, b2;
  betas(mu, kappa, rho_t, Tt, step, c, &b1, &b2);
  const double k_f = -(rho_t * Tt) / 2.0, k_g = -(rho_t * Tt * (2.0 * c->cv - 3.0)) / 2.0;
  int nlinks = 0;
  for (int i = 1; i < NQ; i++) {
    if (state[(int64_t)O39_OPP[i] * n + k] > -1001) continue;
    const int ci[3] = {O39_CX[i], O39_CY[i], O39_CZ[i]};
    const double link = sqrt((double)(ci[0] * ci[0] + ci[1] * ci[1] + ci[2] * ci[2]));
    const double wd = fabs((double)state[(int64_t)i * n + k] + 1001.0) / 1.0e9 * link;
    double du[3][3], dT[3];
    for (int ax = 0; ax < 3; ax++) {
      const int up_ok = state[(int64_t)(4 + ax) * n + k] >= 0;
      const int down_ok = state[(int64_t)(1 + ax) * n + k] >= 0;
      for (int e = 0; e < 3; e++) {
        if (ci[ax] > 0) du[ax][e] = (ut[e] - 0.0) / (1.0 + wd);
        else if (ci[ax] < 0) du[ax][e] = (0.0 - ut[e]) / (1.0 + wd);
        else du[ax][e] = d_local(su[0][e], su[4 + ax][e], su[1 + ax][e], up_ok, down_ok);
      }
      dT[ax] = ci[ax] != 0 ? 0.0 : d_local(sT[0], sT[4 + ax], sT[1 + ax], up_ok, down_ok);
    }
    double F, G;
    o39p_wall_neq(i, Tt, ut, du, dT, b1, b2, c->cv, &F, &G);
    fw[i] = feq[i] + eps * k_f * F;
    gw[i] = c->lead_g * Tt * feq[i] + eps * k_g * G;
    nlinks++;
  }
  return nlinks;
}

/* damp_it.f95:64-123 with a plain-lattice link table: a negative
 * post-collision population takes the mean of its positive neighbours along
 * +-c_i where the node's link codes of i / opp(i) are >= 0 (the rest
 * direction: the node's code), the mean of one with |f|, or |f| */
static int64_t damp_walls_plain(int nx, int ny, int nz, const int32_t fc[6], const int32_t *state,
                                double *fo, int jacobi) {
  const int64_t n = (int64_t)nx * ny * nz;
  double *src = fo;
  if (jacobi) {
    src = (double *)malloc(sizeof(double) * (size_t)(n * NQ));
    memcpy(src, fo, sizeof(double) * (size_t)(n * NQ));
  }
  int64_t nrep = 0;
  for (int i = 0; i < NQ; i++) {
    for (int64_t k = 0; k < n; k++) {
      const double v = src[(int64_t)i * n + k];
      if (!(v < 0.0) || state[k] < 0) continue;
      const int64_t a = k % nx, b = (k / nx) % ny, cc = k / ((int64_t)nx * ny);
      const int64_t kp = pos_index(a + O39_CX[i], b + O39_CY[i], cc + O39_CZ[i], nx, ny, nz, fc);
      const int64_t km = pos_index(a - O39_CX[i], b - O39_CY[i], cc - O39_CZ[i], nx, ny, nz, fc);
      const int pok = kp >= 0 && state[(int64_t)i * n + k] >= 0;
      const int mok = km >= 0 && state[(int64_t)O39_OPP[i] * n + k] >= 0;
      const double fp = pok ? src[(int64_t)i * n + kp] : -1.0;
      const double fm = mok ? src[(int64_t)i * n + km] : -1.0;
      double r;
      if (pok && mok) {
        if (fp > 0.0 && fm > 0.0) r = (fp + fm) / 2.0;
        else if (fp > 0.0 && fm < 0.0) r = (fp + fabs(v)) / 2;
        else if (fp < 0.0 && fm > 0.0) r = (fm + fabs(v)) / 2;
        else r = fabs(v);
      } else if (pok) {
        r = fp > 0.0 ? (fp + fabs(v)) / 2 : fabs(v);
      } else if (mok) {
        r = fm > 0.0 ? (fm + fabs(v)) / 2 : fabs(v);
      } else {
        r = fabs(v);
      }
      fo[(int64_t)i * n + k] = r;
      nrep++;
    }
  }
  if (jacobi) free(src);
  return nrep;
}

/* damp_it.f95:125-177 with the link table (shifted lattice) */
static int64_t damp_walls(int nx, int ny, int nz, const int32_t fc[6], const int32_t *state,
                          double *fo, int jacobi) {
  int cx[NQ], cy[NQ], cz[NQ];
  o39_velocities(1, cx, cy, cz);
  const int64_t n = (int64_t)nx * ny * nz;
  double *src = fo;
  if (jacobi) {
    src = (double *)malloc(sizeof(double) * (size_t)(n * NQ));
    memcpy(src, fo, sizeof(double) * (size_t)(n * NQ));
  }
  int64_t nrep = 0;
  for (int i = 0; i < NQ; i++) {
    for (int64_t k = 0; k < n; k++) {
      const double v = src[(int64_t)i * n + k];
      if (!(v < 0.0) || state[k] < 0) continue;
      const int64_t a = k % nx, b = (k / nx) % ny, cc = k / ((int64_t)nx * ny);
      const int64_t kp = pos_index(a + cx[i], b + cy[i], cc + cz[i], nx, ny, nz, fc);
      const int64_t km = pos_index(a - cx[i], b - cy[i], cc - cz[i], nx, ny, nz, fc);
      /* the rest direction has no link code: its own node code stands in */
      const int32_t link = i == 0 ? state[k] : state[(int64_t)i * n + k];
      const int pok = kp >= 0 && link >= -500;
      const int mok = km >= 0 && state[km] >= -500;
      const double fp = pok ? src[(int64_t)i * n + kp] : -1.0;
      const double fm = mok ? src[(int64_t)i * n + km] : -1.0;
      double r;
      if (pok && mok) {
        if (fp > 0.0 && fm > 0.0) r = (fp + fm) / 2.0;
        else if (fp > 0.0 && fm < 0.0) r = (fp + fabs(v)) / 2;
        else if (fp < 0.0 && fm > 0.0) r = (fm + fabs(v)) / 2;
        else r = fabs(v);
      } else if (pok) {
        r = fp > 0.0 ? (fp + fabs(v)) / 2 : fabs(v);
      } else if (mok) {
        r = fm > 0.0 ? (fm + fabs(v)) / 2 : fabs(v);
      } else {
        continue;
      }
      fo[(int64_t)i * n + k] = r;
      nrep++;
    }
  }
  if (jacobi) free(src);
  return nrep;
}

void o39_step_walls(int nx, int ny, int nz, const int32_t fc[6], const int32_t *state,
                    const double *fs_f, const double *fs_g, const int32_t *face_row,
                    const double *bc_tab, double *f, double *g, double *lm,
                    double *alpha_out, const o39_consts *c, int64_t step0, int64_t lvl_step0,
                    int nsteps, int jacobi, o39_stats *st, int64_t *n_wall_links,
                    int64_t *n_wall_repair) {
  int cx[NQ], cy[NQ], cz[NQ];
  o39_velocities(c->shifted, cx, cy, cz);
  const int64_t n = (int64_t)nx * ny * nz;
  double *fo = (double *)malloc(sizeof(double) * (size_t)(n * NQ));
  double *go = (double *)malloc(sizeof(double) * (size_t)(n * NQ));
  memset(st, 0, sizeof(*st));
  int64_t nwl = 0, nwr = 0;
  for (int s = 0; s < nsteps; s++) {
    const int64_t step = step0 + s, lvl_step = lvl_step0 + s;
    int64_t it_sum = 0, noconv = 0, a2 = 0, alo = 0, ahi = 0, nan = 0, kerr = 0, nodes = 0;
#pragma omp parallel for schedule(dynamic, 64) \
    reduction(+ : it_sum, noconv, a2, alo, ahi, nan, kerr, nodes, nwl, nwr)
    for (int64_t k = 0; k < n; k++) {
      if (state[k] < 0) { /* solid node: unchanged */
        for (int i = 0; i < NQ; i++) {
          fo[(int64_t)i * n + k] = f[(int64_t)i * n + k];
          go[(int64_t)i * n + k] = g[(int64_t)i * n + k];
        }
        if (alpha_out) alpha_out[k] = 2.0;
        continue;
      }
      const int64_t a = k % nx, b = (k / nx) % ny, cc = k / ((int64_t)nx * ny);
      double fi[NQ], gi[NQ], l[NM], e[NQ], u[3], P[6], rho, T;
      int any_wall = 0;
      for (int i = 0; i < NQ; i++) {
        int64_t src;
        int dir;
        const int r = o39_source_bc(a, b, cc, i, nx, ny, nz, fc, face_row, c->shifted, &src, &dir);
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
        if (i > 0 && state[(int64_t)i * n + k] <= -1001) any_wall = 1;
      }
      if (any_wall && c->shifted) {
        double fw[NQ], gw[NQ];
        int nc = 0, ke = 0;
        nwr += o39_wall_rebuild(nx, ny, nz, fc, state, f, g, lm, k, fi, c, step, lvl_step, fw, gw,
                                &nc, &ke);
        noconv += nc;
        kerr += ke;
        for (int i = 1; i < NQ; i++)
          if (state[(int64_t)i * n + k] <= -1001) {
            fi[i] = fw[i];
            gi[i] = gw[i];
            nwl++;
          }
      } else if (any_wall) {
        /* plain lattice: direction i is rebuilt where link opp(i) is a wall link */
        double fw[NQ], gw[NQ];
        int nc = 0, ke = 0;
        nwl += o39p_wall_rebuild(nx, ny, nz, fc, state, f, g, lm, k, fi, c, step, fw, gw, &nc,
                                 &ke);
        noconv += nc;
        kerr += ke;
        for (int i = 1; i < NQ; i++)
          if (state[(int64_t)O39_OPP[i] * n + k] <= -1001) {
            fi[i] = fw[i];
            gi[i] = gw[i];
          }
      }
      o39_moments(fi, gi, c->shifted, &rho, u, &T, P, c->cv);
      for (int j = 0; j < NM; j++) l[j] = lm[(int64_t)j * n + k];
      int it = o39_lm_solve(rho, u, T, P, l, e, c->shifted, 0, c->maxit);
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
      nodes++;
    }
    st->nodes += nodes;
    st->lm_iters += it_sum;
    st->lm_noconv += noconv;
    st->n_alpha2 += a2;
    st->n_alpha_lo += alo;
    st->n_alpha_hi += ahi;
    st->n_nan += nan;
    st->n_kappa_err += kerr;
    st->n_repair += c->shifted ? damp_walls(nx, ny, nz, fc, state, fo, jacobi)
                               : damp_walls_plain(nx, ny, nz, fc, state, fo, jacobi);
    memcpy(f, fo, sizeof(double) * (size_t)(n * NQ));
    memcpy(g, go, sizeof(double) * (size_t)(n * NQ));
  }
  free(fo);
  free(go);
  if (n_wall_links) *n_wall_links = nwl;
  if (n_wall_repair) *n_wall_repair = nwr;
}
