/* D3Q27 step on a box with boundary links, FP64.  TEST INFRASTRUCTURE ONLY
 * (see uclbs_oracle.h).
 *
 * Per fluid node x the incoming population f_i(x) (velocity c_i, arriving
 * from x - c_i) is chosen by the code s of the link from x towards x - c_i,
 * state[opp(i)][x], following the dispatch of looping/timestepping/
 * streaming.f95:365-1133 seen from the receiving node:
 *   s >= 0            f_i = fout_i(x - c_i)                 (:365-368)
 *   s = -10           f_i = fout_opp(i)(x), halfway bounce-back.  The
 *                     reference's bounceback.f95:13-19 swaps fout(i) and
 *                     fout(opp(i)) in place inside a direction-outer push
 *                     loop, which does not return the wall-bound population
 *                     to its node; this is the standard rule it aims at.
 *   -199..-100        f_i = inlet equilibrium j, direction i (inlet_bc.f95:21-53,
 *                     rho_v_bc.f95:23-29: fi(x,opp(k)) = inl_fieq(opp(k)))
 *   -299..-200        f_i = outlet equilibrium j (outlet_bc.f95:23-58)
 *   -1000             f_i = freestream equilibrium (freestream_edge_bc.f95:18-19)
 *   <= -1001          interpolated bounce-back, below
 *   other (-666 ...)  f_i = fout_opp(i)(x)
 * then the node collides as in o_d3q27.c.  Solid nodes keep their values.
 *
 * Interpolated bounce-back (the reference's structure, interpolative_bc/,
 * restated for the isothermal D3Q27 lattice, which the reference lacks):
 *   q_k = |s_k + 1001| / 1e9 for every wall link k of x (target_finder.f95:40);
 *   target velocity u_t = mean over wall links of q_k u(x - c_k) / (q_k + 1),
 *     u(x) when x - c_k is not a fluid link (target_finder.f95:36-83);
 *   target density rho_t = sum over k of f_k(x), with f_k replaced by
 *     f_opp(k)(x) when the link towards x - c_k is a wall (target_finder.f95
 *     :56-60); fout_opp(k)(x) if both sides are walls;
 *   f_i(x) = feq_opp(i)(rho_t, u_t) + f_neq_i, the equilibrium of the
 *     outgoing direction as in interpolative_bounceback_bc.f95:2917 with
 *     streaming.f95:1132, and the Grad/Chapman-Enskog non-equilibrium
 *     f_neq_i = -3 tau rho_t w_i (c_ia c_ib - delta_ab/3) du_b/dx_a,
 *     tau = 1/(2 beta);
 *   du/dx_a from u(x + e_a) and u(x - e_a), a wall side replaced by u = 0
 *     at distance q, a non-fluid, non-wall side by u_t at distance 1
 *     (derivatives.f95:753, first_order_vel_cond_unknown);
 *   a negative result is replaced by its magnitude and counted
 *     (interpolative_bounceback_bc.f95:2969-2973).
 * u(y) are the moments of fout at y (the collision conserves them).
 */
#include "uclbs_oracle.h"
#include <math.h>
#include <stdlib.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static inline int64_t idx(const o_box *b, int x, int y, int z) {
  return x + (int64_t)b->nx * (y + (int64_t)b->ny * z);
}
static inline int wrapi(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }
static inline int is_wall(int32_t s) { return s <= -1001; }

static void vel_at(const o_box *b, const double *f, int64_t k, double u[3]) {
  const int64_t n = (int64_t)b->nx * b->ny * b->nz;
  double fl[27], r;
  for (int i = 0; i < 27; i++) fl[i] = f[(int64_t)i * n + k];
  o27_moments(fl, &r, u);
}

void o27_step_box(const o_box *bx, const int32_t *state, const double *fin, double *fout,
                  const o27_box_params *p, double *alpha_out, o_stats *st, o_bc_stats *bst) {
  const int64_t n = (int64_t)bx->nx * bx->ny * bx->nz;
  int64_t it = 0, nb = 0, nn = 0, nt = 0, nc = 0, nibb = 0, nrep = 0, nodes = 0;
  const double tau = 1.0 / (2.0 * p->beta);
#ifdef _OPENMP
  if (p->nthreads > 0) omp_set_num_threads(p->nthreads);
#pragma omp parallel for collapse(2) schedule(static) \
    reduction(+ : it, nb, nn, nt, nc, nibb, nrep, nodes)
#endif
  for (int z = 0; z < bx->nz; z++)
    for (int y = 0; y < bx->ny; y++)
      for (int x = 0; x < bx->nx; x++) {
        const int64_t k = idx(bx, x, y, z);
        if (state[k] < 0) { /* solid / null node: unchanged */
          for (int i = 0; i < 27; i++) fout[(int64_t)i * n + k] = fin[(int64_t)i * n + k];
          if (alpha_out) alpha_out[k] = 2.0;
          continue;
        }
        double f[27];
        int unknown[27] = {0};
        int any_wall = 0;
        for (int i = 0; i < 27; i++) {
          if (i == 0) { f[0] = fin[k]; continue; }
          const int32_t s = state[(int64_t)O27_OPP[i] * n + k];
          if (s >= 0) {
            int xs = wrapi(x - O27_CX[i], bx->nx), ys = wrapi(y - O27_CY[i], bx->ny),
                zs = wrapi(z - O27_CZ[i], bx->nz);
            f[i] = fin[(int64_t)i * n + idx(bx, xs, ys, zs)];
          } else if (s == -10) {
            f[i] = fin[(int64_t)O27_OPP[i] * n + k];
          } else if (s <= -100 && s >= -199 && -s - 100 <= p->n_in) {
            f[i] = p->feq_in[(size_t)(-s - 101) * 27 + i];
          } else if (s <= -200 && s >= -299 && -s - 200 <= p->n_out) {
            f[i] = p->feq_out[(size_t)(-s - 201) * 27 + i];
          } else if (s == -1000 && p->feq_fs) {
            f[i] = p->feq_fs[i];
          } else if (is_wall(s)) {
            unknown[i] = 1;
            any_wall = 1;
            f[i] = 0.0;
          } else {
            f[i] = fin[(int64_t)O27_OPP[i] * n + k];
          }
        }
        if (any_wall) {
          /* target state */
          double ut[3] = {0, 0, 0}, ux[3];
          int cnt = 0;
          vel_at(bx, fin, k, ux);
          for (int kk = 1; kk < 27; kk++) {
            const int32_t s = state[(int64_t)kk * n + k];
            if (!is_wall(s)) continue;
            double q = fabs((double)s + 1001.0) / 1.0e9;
            double uu[3];
            if (state[(int64_t)O27_OPP[kk] * n + k] >= 0) {
              int xs = wrapi(x - O27_CX[kk], bx->nx), ys = wrapi(y - O27_CY[kk], bx->ny),
                  zs = wrapi(z - O27_CZ[kk], bx->nz);
              vel_at(bx, fin, idx(bx, xs, ys, zs), uu);
            } else {
              uu[0] = ux[0]; uu[1] = ux[1]; uu[2] = ux[2];
            }
            for (int d = 0; d < 3; d++) ut[d] += q * uu[d] / (q + 1.0);
            cnt++;
          }
          for (int d = 0; d < 3; d++) ut[d] /= cnt;
          double rt = 0.0;
          for (int kk = 0; kk < 27; kk++) {
            if (!unknown[kk]) rt += f[kk];
            else if (!unknown[O27_OPP[kk]]) rt += f[O27_OPP[kk]];
            else rt += fin[(int64_t)O27_OPP[kk] * n + k];
          }
          /* velocity gradient G[a][b] = du_b/dx_a */
          double G[3][3];
          for (int a = 0; a < 3; a++) {
            const int dp = 1 + a, dm = 4 + a; /* +e_a, -e_a in the D3Q27 order */
            double hp, hm, up[3], um[3];
            const int32_t sp = state[(int64_t)dp * n + k], sm = state[(int64_t)dm * n + k];
            int ex[3] = {0, 0, 0};
            ex[a] = 1;
            if (sp >= 0) {
              hp = 1.0;
              vel_at(bx, fin, idx(bx, wrapi(x + ex[0], bx->nx), wrapi(y + ex[1], bx->ny),
                                  wrapi(z + ex[2], bx->nz)), up);
            } else if (is_wall(sp)) {
              hp = fabs((double)sp + 1001.0) / 1.0e9;
              up[0] = up[1] = up[2] = 0.0;
            } else {
              hp = 1.0;
              up[0] = ut[0]; up[1] = ut[1]; up[2] = ut[2];
            }
            if (sm >= 0) {
              hm = 1.0;
              vel_at(bx, fin, idx(bx, wrapi(x - ex[0], bx->nx), wrapi(y - ex[1], bx->ny),
                                  wrapi(z - ex[2], bx->nz)), um);
            } else if (is_wall(sm)) {
              hm = fabs((double)sm + 1001.0) / 1.0e9;
              um[0] = um[1] = um[2] = 0.0;
            } else {
              hm = 1.0;
              um[0] = ut[0]; um[1] = ut[1]; um[2] = ut[2];
            }
            for (int b2 = 0; b2 < 3; b2++) G[a][b2] = (up[b2] - um[b2]) / (hp + hm);
          }
          double feq_t[27];
          o27_feq(rt, ut[0], ut[1], ut[2], feq_t);
          for (int i = 1; i < 27; i++) {
            if (!unknown[i]) continue;
            const int c[3] = {O27_CX[i], O27_CY[i], O27_CZ[i]};
            double qs = 0.0;
            for (int a = 0; a < 3; a++)
              for (int b2 = 0; b2 < 3; b2++)
                qs += ((double)(c[a] * c[b2]) - (a == b2 ? 1.0 / 3.0 : 0.0)) * G[a][b2];
            double v = feq_t[O27_OPP[i]] - 3.0 * tau * rt * O27_W[i] * qs;
            if (v < 0.0) { v = -v; nrep++; }
            f[i] = v;
            nibb++;
          }
        }
        double fo[27], rho, u[3];
        o_alpha_result ar =
            o27_collide_node(f, p->beta, 0.0, p->tol, p->maxit, p->alpha_mode, fo, &rho, u);
        for (int i = 0; i < 27; i++) fout[(int64_t)i * n + k] = fo[i];
        if (alpha_out) alpha_out[k] = ar.alpha;
        nodes++;
        it += ar.iters;
        nb += ar.status == O_ALPHA_BOUND;
        nn += ar.status == O_ALPHA_NONPOS;
        nt += ar.status == O_ALPHA_TRIVIAL;
        nc += ar.status == O_ALPHA_NOCONV;
      }
  if (st) {
    st->iters += it;
    st->nodes += nodes;
    st->n_bound += nb;
    st->n_nonpos += nn;
    st->n_trivial += nt;
    st->n_noconv += nc;
  }
  if (bst) {
    bst->n_ibb_links += nibb;
    bst->n_repair += nrep;
  }
}
