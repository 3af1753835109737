/* Link / neighbour table construction, FP64.  TEST INFRASTRUCTURE ONLY (see uclbs_oracle.h).
 *
 * Restates the reference's table builder for one nodal box (level 0):
 *   node fluid/solid:  lbgrid/fluid_or_solid.f95:86-442 -- a ray from the
 *     node away from the grid centre (utilities/ray_casting_away.f95), length
 *     100 x the largest domain extent, tested against every triangle with
 *     utilities/box_test.f95, the plane intersection of fluid_or_solid.f95
 *     :176-196 and the sign test of utilities/tri_volume.f95; odd crossing
 *     parity = solid (external flow: solid_target = 1, :40-46); an odd number
 *     of edge hits with an even number of crossings counts as one crossing
 *     (:393-399).
 *   wall links:  lbgrid/inter_bb_setup.f95:44-246 -- the segment node ->
 *     node + c_i dx, first triangle in index order that it crosses, code
 *     -1001 - floor(|crossing - node| / (|c_i| dx) * 1e9)
 *     (linking/modules/linkwise.f95:45-57); no crossing -> the level (fluid).
 *   domain faces:  lbgrid/state_directions.f95:194-684 -- a link leaving
 *     the box takes its face's code, faces tested in the order x-, y-, z-,
 *     x+, y+, z+; periodic faces wrap instead.
 *   interior links:  state_directions.f95:687-716 -- the neighbour's node
 *     code; solid nodes get -666 on every link (:87).
 * Ties (ray through an edge or vertex, ray in a triangle's plane, node on a
 * surface) make the reference re-cast with a time-seeded random ray
 * (utilities/ray_casting.f95), which no other run can reproduce.  Here such a
 * node is re-cast along fixed directions instead, and counted.
 * Arithmetic follows the reference expression order; built with
 * -ffp-contract=off so no fused multiply-add changes a rounding.
 */
#include "uclbs_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define PLUS_EPS 1.24e-16   /* general/modules/precise.f95:13 */
#define MINUS_EPS (-1.24e-16)
#define VSMALL 1.0e-14      /* precise.f95:16 */

/* utilities/box_test.f95:20-79: 1 if both segment ends lie beyond one face */
static int box_test(const double p1[3], const double p2[3], const double bb[6]) {
  /* bb = xmax, xmin, ymax, ymin, zmax, zmin */
  for (int d = 0; d < 3; d++) {
    double mx = bb[2 * d], mn = bb[2 * d + 1];
    if (p1[d] >= mx + PLUS_EPS && p2[d] >= mx + PLUS_EPS) return 1;
    if (p1[d] <= mn + MINUS_EPS && p2[d] <= mn + MINUS_EPS) return 1;
  }
  return 0;
}

/* utilities/tri_volume.f95:25-108 */
static int tri_volume(const double xc[3], const double e[3], const double *t) {
  double x1 = t[0] - e[0], x2 = t[3] - e[0], x3 = t[6] - e[0];
  double y1 = t[1] - e[1], y2 = t[4] - e[1], y3 = t[7] - e[1];
  double z1 = t[2] - e[2], z2 = t[5] - e[2], z3 = t[8] - e[2];
  double px = xc[0] - e[0], py = xc[1] - e[1], pz = xc[2] - e[2];
  double v[3];
  v[0] = px * (y1 * z2 - z1 * y2) + py * (z1 * x2 - x1 * z2) + pz * (x1 * y2 - y1 * x2);
  v[1] = -px * (y1 * z3 - z1 * y3) - py * (z1 * x3 - x1 * z3) - pz * (x1 * y3 - y1 * x3);
  v[2] = px * (y2 * z3 - z2 * y3) + py * (z2 * x3 - x2 * z3) + pz * (x2 * y3 - y2 * x3);
  int zc = 0;
  for (int i = 0; i < 3; i++)
    if (v[i] >= MINUS_EPS && v[i] <= PLUS_EPS) zc++;
  if (v[0] > 0.0 && v[1] > 0.0 && v[2] > 0.0) return 1;
  if (v[0] < 0.0 && v[1] < 0.0 && v[2] < 0.0) return 1;
  if (zc == 1) return -1;
  if (zc == 2) return -2;
  if (zc == 3) return -3;
  return 0;
}

/* Triangle t: 9 vertex coordinates, normal n, bounding box bb. */
static void tri_bbox(const double *t, double bb[6]) {
  for (int d = 0; d < 3; d++) {
    double a = t[d], b = t[3 + d], c = t[6 + d];
    double mx = a > b ? a : b; mx = mx > c ? mx : c;
    double mn = a < b ? a : b; mn = mn < c ? mn : c;
    bb[2 * d] = mx;
    bb[2 * d + 1] = mn;
  }
}

/* Parity cast of one ray.  Returns 0 (even), 1 (odd) or -1 (a tie). */
static int cast_parity(const double e1[3], const double e2[3], int ntri, const double *tri,
                       const double *nrm, const double *tbb) {
  int cross = 0, edge = 0;
  for (int i = 0; i < ntri; i++) {
    const double *t = tri + 9 * (size_t)i, *n = nrm + 3 * (size_t)i;
    if (box_test(e1, e2, tbb + 6 * (size_t)i)) continue;
    double dot = n[0] * t[0] + n[1] * t[1] + n[2] * t[2];
    double upper = dot - (e1[0] * n[0] + e1[1] * n[1] + e1[2] * n[2]);
    double lower = (e2[0] - e1[0]) * n[0] + (e2[1] - e1[1]) * n[1] + (e2[2] - e1[2]) * n[2];
    if (lower >= MINUS_EPS && lower <= PLUS_EPS) return -1; /* ray in / parallel to the plane */
    double pc = upper / lower;
    if (pc > 0.0 && pc < 1.0) {
      double xc[3] = {e1[0] + pc * (e2[0] - e1[0]), e1[1] + pc * (e2[1] - e1[1]),
                      e1[2] + pc * (e2[2] - e1[2])};
      if (upper >= lower + MINUS_EPS && upper <= lower + PLUS_EPS) return -1;
      int v = tri_volume(xc, e1, t);
      if (v == 1) cross++;
      else if (v == -1) edge++;
      else if (v <= -2) return -1;
    } else if (upper >= -VSMALL && upper <= VSMALL) {
      return -1; /* node on the triangle's plane */
    }
  }
  if (edge >= 2) return -1; /* corner casting in the reference */
  if ((edge % 2) == 1 && (cross % 2) == 0) { edge = 0; cross++; }
  edge /= 2;
  return (cross + edge) % 2;
}

/* fixed re-cast directions for ties (unit vectors, irrational-ish components) */
static const double RECAST[3][3] = {{0.5773502691896258, 0.6311775151936213, 0.5179036655461733},
                                    {-0.7071067811865476, 0.3826834323650898, 0.5946035575013605},
                                    {0.2672612419124244, -0.5345224838248488, 0.8017837257372732}};

int o_classify_node(const o_box *bx, int ntri, const double *tri, const double *nrm,
                    const double *tbb, const double gbb[6], int a, int b, int c, int *tie) {
  double e1[3] = {bx->origin[0] + (double)a * bx->dx, bx->origin[1] + (double)b * bx->dx,
                  bx->origin[2] + (double)c * bx->dx};
  double cen[3], ext = 0.0;
  for (int d = 0; d < 3; d++) {
    int nd = d == 0 ? bx->nx : (d == 1 ? bx->ny : bx->nz);
    double lo = bx->origin[d], hi = bx->origin[d] + (double)(nd - 1) * bx->dx;
    cen[d] = (lo + hi) / 2.0;
    if (hi - lo > ext) ext = hi - lo;
  }
  double L = fabs(100.0 * ext);
  double dl[3] = {e1[0] - cen[0], e1[1] - cen[1], e1[2] - cen[2]};
  double dist = sqrt(dl[0] * dl[0] + dl[1] * dl[1] + dl[2] * dl[2]);
  double e2[3];
  *tie = 0;
  for (int attempt = 0; attempt < 4; attempt++) {
    if (attempt == 0 && dist > 0.0) {
      for (int d = 0; d < 3; d++) e2[d] = e1[d] + L * dl[d] / dist;
    } else {
      const double *r = RECAST[attempt == 0 ? 0 : attempt - 1];
      for (int d = 0; d < 3; d++) e2[d] = e1[d] + L * r[d];
      if (attempt == 0) *tie = 1;
    }
    if (box_test(e1, e2, gbb)) return 0; /* misses the geometry: fluid */
    int p = cast_parity(e1, e2, ntri, tri, nrm, tbb);
    if (p >= 0) return p; /* 1 = solid */
    *tie = 1;
  }
  return 0;
}

/* linkwise.f95:45-57: code of a crossing xc on a link of length link from e1 */
int32_t o_link_code(const double e1[3], const double xc[3], double link) {
  double d = sqrt((e1[0] - xc[0]) * (e1[0] - xc[0]) + (e1[1] - xc[1]) * (e1[1] - xc[1]) +
                  (e1[2] - xc[2]) * (e1[2] - xc[2]));
  return (int32_t)(-1001 - (int64_t)floor(d / link * 1e9));
}

int o_tri_volume(const double xc[3], const double e[3], const double t[9]) {
  return tri_volume(xc, e, t);
}

int o_box_test(const double p1[3], const double p2[3], const double bb[6]) {
  return box_test(p1, p2, bb);
}

/* inter_bb_setup for the segment node -> node + c dx; returns the link code */
int32_t o_wall_code(const o_box *bx, int ntri, const double *tri, const double *nrm,
                    const double *tbb, int a, int b, int c, int cx, int cy, int cz, int *tie) {
  double e1[3] = {bx->origin[0] + (double)a * bx->dx, bx->origin[1] + (double)b * bx->dx,
                  bx->origin[2] + (double)c * bx->dx};
  double e2[3] = {e1[0] + (double)cx * bx->dx, e1[1] + (double)cy * bx->dx,
                  e1[2] + (double)cz * bx->dx};
  double link = sqrt((double)(cx * cx + cy * cy + cz * cz)) * bx->dx; /* link_info.f95:29 */
  *tie = 0;
  for (int i = 0; i < ntri; i++) {
    const double *t = tri + 9 * (size_t)i, *n = nrm + 3 * (size_t)i;
    if (box_test(e1, e2, tbb + 6 * (size_t)i)) continue;
    double dot = n[0] * t[0] + n[1] * t[1] + n[2] * t[2];
    double upper = dot - (e1[0] * n[0] + e1[1] * n[1] + e1[2] * n[2]);
    double lower = (e2[0] - e1[0]) * n[0] + (e2[1] - e1[1]) * n[1] + (e2[2] - e1[2]) * n[2];
    if (lower >= MINUS_EPS && lower <= PLUS_EPS) continue;
    double pc = upper / lower;
    if (pc > 0.0 && pc < 1.0 + PLUS_EPS) {
      double xc[3] = {e1[0] + pc * (e2[0] - e1[0]), e1[1] + pc * (e2[1] - e1[1]),
                      e1[2] + pc * (e2[2] - e1[2])};
      int hit;
      if (upper >= lower - VSMALL && upper <= lower + VSMALL) { hit = 1; *tie = 1; }
      else hit = tri_volume(xc, e1, t) != 0;
      if (hit) {
        return o_link_code(e1, xc, link);
      }
    } else if (upper >= -VSMALL && upper <= VSMALL) {
      *tie = 1; /* the reference nudges the node off the surface and retries */
    }
  }
  return 0; /* no crossing: the level of the node (fluid link) */
}

static int face_of(const o_box *bx, int a, int b, int c, int cx, int cy, int cz) {
  if (a + cx < 0) return 0;
  if (b + cy < 0) return 1;
  if (c + cz < 0) return 2;
  if (a + cx > bx->nx - 1) return 3;
  if (b + cy > bx->ny - 1) return 4;
  if (c + cz > bx->nz - 1) return 5;
  return -1;
}

int64_t o_build_links27(const o_box *bx, int ntri, const double *tri, const double *nrm,
                        int32_t *state, int64_t *n_tie) {
  const int64_t n = (int64_t)bx->nx * bx->ny * bx->nz;
  double *tbb = (double *)malloc(sizeof(double) * 6 * (size_t)(ntri > 0 ? ntri : 1));
  double gbb[6] = {-1e300, 1e300, -1e300, 1e300, -1e300, 1e300};
  for (int i = 0; i < ntri; i++) {
    tri_bbox(tri + 9 * (size_t)i, tbb + 6 * (size_t)i);
    for (int d = 0; d < 3; d++) {
      if (tbb[6 * i + 2 * d] > gbb[2 * d]) gbb[2 * d] = tbb[6 * i + 2 * d];
      if (tbb[6 * i + 2 * d + 1] < gbb[2 * d + 1]) gbb[2 * d + 1] = tbb[6 * i + 2 * d + 1];
    }
  }
  int64_t ties = 0, nsolid = 0;
  /* pass 1: node codes (component 0) */
#ifdef _OPENMP
#pragma omp parallel for collapse(2) schedule(dynamic, 4) reduction(+ : ties, nsolid)
#endif
  for (int c = 0; c < bx->nz; c++)
    for (int b = 0; b < bx->ny; b++)
      for (int a = 0; a < bx->nx; a++) {
        int64_t k = a + (int64_t)bx->nx * (b + (int64_t)bx->ny * c);
        int tie = 0;
        int solid = ntri > 0 ? o_classify_node(bx, ntri, tri, nrm, tbb, gbb, a, b, c, &tie) : 0;
        state[k] = solid ? -1001 : 0;
        ties += tie;
        nsolid += solid;
      }
  /* pass 2: link codes */
#ifdef _OPENMP
#pragma omp parallel for collapse(2) schedule(dynamic, 4) reduction(+ : ties)
#endif
  for (int c = 0; c < bx->nz; c++)
    for (int b = 0; b < bx->ny; b++)
      for (int a = 0; a < bx->nx; a++) {
        int64_t k = a + (int64_t)bx->nx * (b + (int64_t)bx->ny * c);
        int32_t own = state[k];
        for (int i = 1; i < 27; i++) {
          int32_t code;
          if (own == -1001 || own == -666) {
            code = -666;
          } else {
            int cx = O27_CX[i], cy = O27_CY[i], cz = O27_CZ[i];
            int f = face_of(bx, a, b, c, cx, cy, cz);
            int na = a + cx, nb = b + cy, nc = c + cz;
            int wrap = 1;
            if (f >= 0) {
              /* periodic faces wrap; a link leaving through several faces uses
               * the first non-periodic one in reference order */
              for (int g = 0; g < 6 && wrap; g++) {
                int out = (g == 0 && na < 0) || (g == 1 && nb < 0) || (g == 2 && nc < 0) ||
                          (g == 3 && na > bx->nx - 1) || (g == 4 && nb > bx->ny - 1) ||
                          (g == 5 && nc > bx->nz - 1);
                if (out && bx->face_code[g] != O_FACE_PERIODIC) { wrap = 0; f = g; }
              }
            }
            if (f >= 0 && !wrap) {
              code = bx->face_code[f];
            } else {
              na = (na + bx->nx) % bx->nx;
              nb = (nb + bx->ny) % bx->ny;
              nc = (nc + bx->nz) % bx->nz;
              int32_t nbr = state[na + (int64_t)bx->nx * (nb + (int64_t)bx->ny * nc)];
              if (nbr == -1001) {
                int tie = 0;
                code = o_wall_code(bx, ntri, tri, nrm, tbb, a, b, c, cx, cy, cz, &tie);
                ties += tie;
              } else {
                code = nbr;
              }
            }
          }
          state[(int64_t)i * n + k] = code;
        }
      }
  free(tbb);
  if (n_tie) *n_tie = ties;
  return nsolid;
}

/* D3Q39 table on the plain lattice: lbgrid/state_directions.f95 with the
 * velocities of directions_and_weights.f95:436-444.  Links are set direction
 * by direction (the reference's outer loop), so a long link can read the
 * codes of its shorter links, which are already final:
 *   (+-2,0,0)-type, Fortran 16-21 (:95-125): when the speed-1 link i-14 is a
 *     wall link s, the code is (s + 1000) / 2 (integer division toward zero;
 *     s = -1001 gives 0, a fluid code);
 *   (+-3,0,0)-type, Fortran 34-39 (:127-185): (s1 + 1000) / 3 from the
 *     speed-1 link i-32, else ((s2 + 1000) * 2) / 3 from the speed-2 link
 *     i-18 -- except Fortran 35, which reads its own slot state(...,35)
 *     (:141-146), a value the pass has not written yet (AMReX leaves the
 *     array uninitialised; the table starts at 0 here, as on the device).
 * Every other link: face code (an outflow face, -6, gives 0), neighbour's
 * node code, or the wall code of the segment node -> node + c dx
 * (inter_bb_setup.f95:44-246), as on D3Q27.
 * state must hold 39 x n int32 and is cleared here. */
int64_t o_build_links39(const o_box *bx, int ntri, const double *tri, const double *nrm,
                        int32_t *state, int64_t *n_tie) {
  const int64_t n = (int64_t)bx->nx * bx->ny * bx->nz;
  memset(state, 0, sizeof(int32_t) * (size_t)(39 * n));
  double *tbb = (double *)malloc(sizeof(double) * 6 * (size_t)(ntri > 0 ? ntri : 1));
  double gbb[6] = {-1e300, 1e300, -1e300, 1e300, -1e300, 1e300};
  for (int i = 0; i < ntri; i++) {
    tri_bbox(tri + 9 * (size_t)i, tbb + 6 * (size_t)i);
    for (int d = 0; d < 3; d++) {
      if (tbb[6 * i + 2 * d] > gbb[2 * d]) gbb[2 * d] = tbb[6 * i + 2 * d];
      if (tbb[6 * i + 2 * d + 1] < gbb[2 * d + 1]) gbb[2 * d + 1] = tbb[6 * i + 2 * d + 1];
    }
  }
  int64_t ties = 0, nsolid = 0;
#ifdef _OPENMP
#pragma omp parallel for collapse(2) schedule(dynamic, 4) reduction(+ : ties, nsolid)
#endif
  for (int c = 0; c < bx->nz; c++)
    for (int b = 0; b < bx->ny; b++)
      for (int a = 0; a < bx->nx; a++) {
        int64_t k = a + (int64_t)bx->nx * (b + (int64_t)bx->ny * c);
        int tie = 0;
        int solid = ntri > 0 ? o_classify_node(bx, ntri, tri, nrm, tbb, gbb, a, b, c, &tie) : 0;
        state[k] = solid ? -1001 : 0;
        ties += tie;
        nsolid += solid;
      }
#ifdef _OPENMP
#pragma omp parallel for collapse(2) schedule(dynamic, 4) reduction(+ : ties)
#endif
  for (int c = 0; c < bx->nz; c++)
    for (int b = 0; b < bx->ny; b++)
      for (int a = 0; a < bx->nx; a++) {
        int64_t k = a + (int64_t)bx->nx * (b + (int64_t)bx->ny * c);
        int32_t own = state[k];
        for (int i = 1; i < 39; i++) {      /* Fortran direction i + 1 */
          int32_t code;
          if (own == -1001 || own == -666) {
            state[(int64_t)i * n + k] = -666;
            continue;
          }
          if (i >= 15 && i <= 20) {          /* Fortran 16-21 */
            int32_t s1 = state[(int64_t)(i - 14) * n + k];
            if (s1 < -1000) {
              state[(int64_t)i * n + k] = (s1 + 1000) / 2;
              continue;
            }
          } else if (i >= 33) {              /* Fortran 34-39 */
            int32_t s1 = state[(int64_t)(i - 32) * n + k];
            int32_t s2 = state[(int64_t)(i - 18) * n + k];
            int32_t s2r = (i == 34) ? state[(int64_t)i * n + k] : s2;
            if (s1 < -1000) {
              state[(int64_t)i * n + k] = (s1 + 1000) / 3;
              continue;
            } else if (s2 < -1000) {
              state[(int64_t)i * n + k] = ((s2r + 1000) * 2) / 3;
              continue;
            }
          }
          int cx = O39_CX[i], cy = O39_CY[i], cz = O39_CZ[i];
          int f = face_of(bx, a, b, c, cx, cy, cz);
          int na = a + cx, nb = b + cy, nc = c + cz;
          int wrap = 1;
          if (f >= 0) {
            for (int g = 0; g < 6 && wrap; g++) {
              int out = (g == 0 && na < 0) || (g == 1 && nb < 0) || (g == 2 && nc < 0) ||
                        (g == 3 && na > bx->nx - 1) || (g == 4 && nb > bx->ny - 1) ||
                        (g == 5 && nc > bx->nz - 1);
              if (out && bx->face_code[g] != O_FACE_PERIODIC) { wrap = 0; f = g; }
            }
          }
          if (f >= 0 && !wrap) {
            /* an outflow face (outlet type 6) gives the code of the node
             * outside, a fluid node of level 0 (state_directions_lvl.f95:131) */
            code = bx->face_code[f] == -6 ? 0 : bx->face_code[f];
          } else {
            na = ((na % bx->nx) + bx->nx) % bx->nx;
            nb = ((nb % bx->ny) + bx->ny) % bx->ny;
            nc = ((nc % bx->nz) + bx->nz) % bx->nz;
            int32_t nbr = state[na + (int64_t)bx->nx * (nb + (int64_t)bx->ny * nc)];
            if (nbr == -1001) {
              int tie = 0;
              code = o_wall_code(bx, ntri, tri, nrm, tbb, a, b, c, cx, cy, cz, &tie);
              ties += tie;
            } else {
              code = nbr;
            }
          }
          state[(int64_t)i * n + k] = code;
        }
      }
  free(tbb);
  if (n_tie) *n_tie = ties;
  return nsolid;
}

/* D3Q39 table on the shifted lattice (c_x + 1): lbgrid/
 * shifted_state_directions.f95:47-737.  Every link i of a fluid node x is
 * keyed on its upstream node u = x - c_i (the node the pull reads):
 *   u beyond a non-periodic face (first in the order x-, y-, z-, x+, y+, z+,
 *     :60-329):  the wall code of the segment x -> u when u is a solid node
 *     (inter_bb_setup.f95 with the ray reversed for the shifted lattice,
 *     :37-47), else the face code; an outflow face (outlet type 6) gives the
 *     code of the downstream node x + c_i (:88-89);
 *   u inside (or wrapped across a periodic face):  the wall code when u is
 *     solid, else the downstream node's code (:331-339);
 *   a code of -1001 (a solid downstream node, or a wall segment that crosses
 *     no triangle) becomes the level, 0 (:345).
 * No long link is derived from a shorter one (each is ray-tested), and the
 * shifted rest direction (Fortran 1) keeps the node's own code.  A node
 * outside the box (an upstream beyond a face, a downstream beyond the box)
 * is classified by the same ray cast as the box's nodes; an out-of-box
 * downstream node's code is 0 (fluid) or, solid, -1001 -> 0 as above.
 * state must hold 39 x n int32. */
int64_t o_build_links39s(const o_box *bx, int ntri, const double *tri, const double *nrm,
                         int32_t *state, int64_t *n_tie) {
  const int64_t n = (int64_t)bx->nx * bx->ny * bx->nz;
  memset(state, 0, sizeof(int32_t) * (size_t)(39 * n));
  double *tbb = (double *)malloc(sizeof(double) * 6 * (size_t)(ntri > 0 ? ntri : 1));
  double gbb[6] = {-1e300, 1e300, -1e300, 1e300, -1e300, 1e300};
  for (int i = 0; i < ntri; i++) {
    tri_bbox(tri + 9 * (size_t)i, tbb + 6 * (size_t)i);
    for (int d = 0; d < 3; d++) {
      if (tbb[6 * i + 2 * d] > gbb[2 * d]) gbb[2 * d] = tbb[6 * i + 2 * d];
      if (tbb[6 * i + 2 * d + 1] < gbb[2 * d + 1]) gbb[2 * d + 1] = tbb[6 * i + 2 * d + 1];
    }
  }
  int64_t ties = 0, nsolid = 0;
#ifdef _OPENMP
#pragma omp parallel for collapse(2) schedule(dynamic, 4) reduction(+ : ties, nsolid)
#endif
  for (int c = 0; c < bx->nz; c++)
    for (int b = 0; b < bx->ny; b++)
      for (int a = 0; a < bx->nx; a++) {
        int64_t k = a + (int64_t)bx->nx * (b + (int64_t)bx->ny * c);
        int tie = 0;
        int solid = ntri > 0 ? o_classify_node(bx, ntri, tri, nrm, tbb, gbb, a, b, c, &tie) : 0;
        state[k] = solid ? -1001 : 0;
        ties += tie;
        nsolid += solid;
      }
  const int nd[3] = {bx->nx, bx->ny, bx->nz};
#ifdef _OPENMP
#pragma omp parallel for collapse(2) schedule(dynamic, 4) reduction(+ : ties)
#endif
  for (int c = 0; c < bx->nz; c++)
    for (int b = 0; b < bx->ny; b++)
      for (int a = 0; a < bx->nx; a++) {
        int64_t k = a + (int64_t)bx->nx * (b + (int64_t)bx->ny * c);
        int32_t own = state[k];
        for (int i = 1; i < 39; i++) {
          int32_t code;
          if (own == -1001 || own == -666) {
            state[(int64_t)i * n + k] = -666;
            continue;
          }
          const int cv[3] = {O39_CX[i] + 1, O39_CY[i], O39_CZ[i]};
          const int x[3] = {a, b, c};
          int up[3], dn[3];
          for (int d = 0; d < 3; d++) { up[d] = x[d] - cv[d]; dn[d] = x[d] + cv[d]; }
          int face = -1;
          for (int f = 0; f < 6 && face < 0; f++) {
            int d = f % 3;
            int out = f < 3 ? up[d] < 0 : up[d] > nd[d] - 1;
            if (out && bx->face_code[f] != O_FACE_PERIODIC) face = f;
          }
          for (int d = 0; d < 3; d++) {
            if (up[d] < 0 && bx->face_code[d] == O_FACE_PERIODIC) up[d] += nd[d];
            if (up[d] > nd[d] - 1 && bx->face_code[d + 3] == O_FACE_PERIODIC) up[d] -= nd[d];
            if (dn[d] < 0 && bx->face_code[d] == O_FACE_PERIODIC) dn[d] += nd[d];
            if (dn[d] > nd[d] - 1 && bx->face_code[d + 3] == O_FACE_PERIODIC) dn[d] -= nd[d];
          }
          int up_in = 1, dn_in = 1;
          for (int d = 0; d < 3; d++) {
            up_in = up_in && up[d] >= 0 && up[d] < nd[d];
            dn_in = dn_in && dn[d] >= 0 && dn[d] < nd[d];
          }
          int up_solid;
          if (up_in) {
            up_solid = state[up[0] + (int64_t)bx->nx * (up[1] + (int64_t)bx->ny * up[2])] < -1000;
          } else {
            int tie = 0;
            up_solid = ntri > 0 &&
                       o_classify_node(bx, ntri, tri, nrm, tbb, gbb, up[0], up[1], up[2], &tie);
            ties += tie;
          }
          int32_t down = dn_in ? state[dn[0] + (int64_t)bx->nx * (dn[1] + (int64_t)bx->ny * dn[2])] : 0;
          if (up_solid) {
            int tie = 0;
            code = o_wall_code(bx, ntri, tri, nrm, tbb, a, b, c, -cv[0], -cv[1], -cv[2], &tie);
            ties += tie;
          } else if (face >= 0) {
            code = bx->face_code[face] == -6 ? down : bx->face_code[face];
          } else {
            code = down;
          }
          if (code == -1001) code = 0;
          state[(int64_t)i * n + k] = code;
        }
      }
  free(tbb);
  if (n_tie) *n_tie = ties;
  return nsolid;
}
