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
 *   <= -1001          linear interpolated bounce-back, below
 *   other (-666 ...)  f_i = fout_opp(i)(x)
 * then the node collides as in o_d3q27.c.  Solid nodes keep their values.
 *
 * Wall links (the reference's D3Q27 lattice has no wall rule; its D3Q39
 * interpolated bounce-back lives in interpolative_bc/).  The missing
 * population f_i(x) of a link whose wall lies at q = |s + 1001| / 1e9
 * (target_finder.f95:40, linkwise.f95:52-56) of the way from x towards
 * x - c_i is the linear interpolated bounce-back of Bouzidi, Firdaouss &
 * Lallemand (2001) on the post-collision populations, k = opp(i):
 *   q <  1/2:  f_i(x) = 2q fout_k(x) + (1 - 2q) fout_k(x + c_i)
 *   q >= 1/2:  f_i(x) = fout_k(x) / (2q) + (1 - 1/(2q)) fout_i(x)
 * (halfway bounce-back fout_k(x) where x + c_i is not a fluid node).  Both
 * forms are convex combinations, so no population turns negative.
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
          /* linear interpolated bounce-back (Bouzidi, Firdaouss & Lallemand,
           * Phys. Fluids 13, 3452 (2001)) from the post-collision values fin:
           * k = opp(i) points into the wall, which lies at fraction q of the
           * link from x towards x - c_i */
          for (int i = 1; i < 27; i++) {
            if (!unknown[i]) continue;
            const int kk = O27_OPP[i];
            const double q = fabs((double)state[(int64_t)kk * n + k] + 1001.0) / 1.0e9;
            const double fk = fin[(int64_t)kk * n + k];
            double v;
            if (q >= 0.5) {
              v = fk / (2.0 * q) + (1.0 - 1.0 / (2.0 * q)) * fin[(int64_t)i * n + k];
            } else if (state[(int64_t)i * n + k] >= 0) {
              /* the fluid node behind x, x + c_i */
              const int xs = wrapi(x + O27_CX[i], bx->nx), ys = wrapi(y + O27_CY[i], bx->ny),
                        zs = wrapi(z + O27_CZ[i], bx->nz);
              v = 2.0 * q * fk + (1.0 - 2.0 * q) * fin[(int64_t)kk * n + idx(bx, xs, ys, zs)];
            } else {
              v = fk; /* no fluid node behind x: halfway bounce-back */
            }
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
