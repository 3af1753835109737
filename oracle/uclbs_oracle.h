/*
 * uclbs_oracle.h -- FP64 CPU oracle for the ELBM collide-stream path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP
 * product path is compared against.  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load it; the product library never
 * links or calls it.
 *
 * Every routine restates an algorithm of the UCLBS reference (Fortran,
 * v0.08) or, where the reference has no implementation (D3Q27), the
 * published method the GPU library uses in its place.  Each definition below
 * names the reference file:line it follows.  All arithmetic is FP64, as in
 * the reference (general/modules/precise.f95:12).
 *
 * Population layout used by every grid routine: structure-of-arrays
 * f[q * n_nodes + node], node = x + nx*(y + ny*z)  (x fastest), which is
 * the component-major Fortran order of an AMReX FAB fi(a,b,c,q) without
 * ghost cells.
 */
#ifndef UCLBS_ORACLE_H
#define UCLBS_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/* ---------------- lattices (general/background_calcs/directions_and_weights.f95) */
/* D3Q27 velocity order of linking/modules/linkwise.f95:30-37 (cx_27 .. opp_27),
 * 0-based here.  Weights are the tensor products of {1/6, 2/3, 1/6}. */
extern const int O27_CX[27], O27_CY[27], O27_CZ[27], O27_OPP[27];
extern const double O27_W[27];
/* D3Q39 (directions_and_weights.f95:436-444; opp :498-501; erp linkwise.f95:39-41) */
extern const int O39_CX[39], O39_CY[39], O39_CZ[39], O39_OPP[39], O39_ERP[39];

/* ---------------- Philox4x32-10 (Salmon et al., SC'11) ----------------------- */
void o_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
/* uniform double in [0,1) from (node, component, seed): 53 bits of out[0..1] */
double o_uniform(uint64_t node, uint32_t comp, uint64_t seed);

/* ---------------- D3Q27 entropic LBM (no reference implementation) ------------ */
/* Product-form entropic equilibrium (Ansumali, Karlin & Ottinger 2003). */
void o27_feq(double rho, double ux, double uy, double uz, double feq[27]);
void o27_moments(const double f[27], double *rho, double u[3]);

/* Entropic alpha: nontrivial root of H(f + a*D) = H(f), H = sum f ln(f/w),
 * D = feq - f.  Newton definition: calculate_alpha.f95:416-556 (commented-out
 * in the reference; positivity bracket :424-445, "alpha_max < 2" rule :449-455).
 * Status codes (also used by the GPU path):                                  */
enum { O_ALPHA_NEWTON = 0, O_ALPHA_TRIVIAL = 1, O_ALPHA_BOUND = 2,
       O_ALPHA_NONPOS = 3, O_ALPHA_NOCONV = 4 };
typedef struct { double alpha; int iters; int status; double resid; } o_alpha_result;
/* alpha0 <= 0 means the plain initial guess 2.  tol: |g(a)| <= tol * S2. */
o_alpha_result o27_alpha(const double f[27], const double feq[27], double alpha0,
                         double tol, int maxit);
/* Same root via the power-sum form of g (used to check the GPU's evaluation). */
o_alpha_result o27_alpha_series(const double f[27], const double feq[27], double alpha0,
                                double tol, int maxit, int order);
/* g(a) and g'(a) of the entropy equality in the conditioned form (log path). */
void o27_entropy_g(const double f[27], const double feq[27], double a, double *g, double *gp);
/* Direct H(f) = sum f ln(f/w) (used by tests in long double). */
long double o27_H_ld(const long double f[27]);

/* one node: stream result f -> post-collision fout; returns alpha result */
o_alpha_result o27_collide_node(const double f[27], double beta, double alpha0, double tol,
                                int maxit, int fixed_alpha_mode, double fout[27],
                                double *rho, double u[3]);

/* Periodic fused step on an nx*ny*nz grid: fout_new(x) = collide(pull(fin)).
 * alpha_out, iters_out may be NULL.  Returns the sum of Newton iterations. */
typedef struct {
  double beta;          /* 1/(6 nu + 1) */
  double tol;           /* alpha residual tolerance */
  int maxit;            /* Newton cap */
  int alpha_mode;       /* 0 = entropic Newton (plain), 1 = fixed alpha = 2 (LBGK) */
  int nthreads;         /* OpenMP threads, 0 = default */
} o27_params;
typedef struct { int64_t iters; int64_t nodes; int64_t n_bound; int64_t n_nonpos;
                 int64_t n_trivial; int64_t n_noconv; } o_stats;
void o27_step_periodic(int nx, int ny, int nz, const double *fin, double *fout,
                       const o27_params *p, double *alpha_out, o_stats *st);
/* Taylor-Green vortex initial state: f = feq(1, u), u = U0*(sin x cos y cos z,
 * -cos x sin y cos z, 0) + U(-amp, amp) noise per component from Philox. */
void o27_init_tgv(int nx, int ny, int nz, double u0, double noise_amp, uint64_t seed,
                  double *f);
void o27_macro(int n, const double *f, double *rho, double *ux, double *uy, double *uz);

/* ---------------- link tables (lbgrid/, linking/modules/linkwise.f95) --------- */
/* Face codes of a box: the reference's edge classification (state_directions.f95
 * :194-684): inlet j -> -100-j, outlet j -> -200-j, bounce-back wall -10,
 * freestream -1000; O_FACE_PERIODIC wraps the axis instead. */
#define O_FACE_PERIODIC 2147483647
typedef struct {
  int nx, ny, nz;      /* nodes per axis (nodal box, level 0) */
  double dx;           /* node spacing */
  double origin[3];    /* coordinate of node (0,0,0) (amrex_problo) */
  int32_t face_code[6];/* x-, y-, z-, x+, y+, z+ */
} o_box;
/* state[q*n + node] in the reference encoding (component 0: node code 0 fluid /
 * -1001 solid; q>0: link code).  Triangles: tri[t*9 + vertex*3 + axis],
 * nrm[t*3 + axis].  Returns the number of solid nodes; *n_tie counts ties
 * resolved by a deterministic re-cast. */
int64_t o_build_links27(const o_box *bx, int ntri, const double *tri, const double *nrm,
                        int32_t *state, int64_t *n_tie);
int64_t o_build_links39(const o_box *bx, int ntri, const double *tri, const double *nrm,
                        int32_t *state, int64_t *n_tie);
int64_t o_build_links39s(const o_box *bx, int ntri, const double *tri, const double *nrm,
                         int32_t *state, int64_t *n_tie);
/* geometry helpers of the table builder, exported for checks against the
 * reference's own routines: tri_volume.f95, box_test.f95 (bb = xmax, xmin,
 * ymax, ymin, zmax, zmin), linkwise.f95:45-57 */
int o_tri_volume(const double xc[3], const double e[3], const double t[9]);
int o_box_test(const double p1[3], const double p2[3], const double bb[6]);
int32_t o_link_code(const double e1[3], const double xc[3], double link);
int32_t o_wall_code(const o_box *bx, int ntri, const double *tri, const double *nrm,
                    const double *tbb, int a, int b, int c, int cx, int cy, int cz, int *tie);
int o_classify_node(const o_box *bx, int ntri, const double *tri, const double *nrm,
                    const double *tbb, const double gbb[6], int a, int b, int c, int *tie);

/* ---------------- D3Q27 step on a box with boundary links ---------------------- */
typedef struct {
  double beta, tol;
  int maxit, alpha_mode, nthreads;
  int n_in, n_out;
  const double *feq_in;  /* [n_in][27]: inlet equilibria (inlet_outlet_basis_calcs.f90:113-119) */
  const double *feq_out; /* [n_out][27] */
  const double *feq_fs;  /* [27] freestream equilibrium (freestream_basis_calcs.f90:66-71) */
} o27_box_params;
typedef struct { int64_t n_ibb_links; int64_t n_repair; } o_bc_stats;
void o27_step_box(const o_box *bx, const int32_t *state, const double *fin, double *fout,
                  const o27_box_params *p, double *alpha_out, o_stats *st, o_bc_stats *bst);


/* ---- D3Q39 two-population compressible step (o_d3q39.c) ---------------- */
typedef struct o39_consts {
  int shifted, turbulent, lvl, startup, maxit;
  double gama, R, cv, cp, lead_g, t_ref, p_ref, rho_ref, mu_ref, kappa_ref, sos, mach;
  double molec_weight, dt, eps, init_beta1, init_beta2, beta1_scale, beta2_scale;
} o39_consts;
typedef struct o39_stats {
  int64_t nodes, lm_iters, lm_noconv, n_alpha2, n_alpha_lo, n_alpha_hi, n_nan, n_kappa_err,
      n_repair;
} o39_stats;
void o39_velocities(int shifted, int cx[39], int cy[39], int cz[39]);
void o39_consts_make(o39_consts *c, double gama, double temperature_k, double pressure_pa,
                     double mach, double dx_m, int startup, int shifted, int turbulent, int lvl);
double o39_sutherland(double T, const o39_consts *c, double *dim_mu, double *dim_temp);
int o39_therm_cond(double rho, double T, const o39_consts *c, double *kappa);
void o39_feq(const double lm[13], double rho, int shifted, double feq[39]);
double o39_lu_step(const double *J, const double *mom, double *lm, int ref_mode);
double o39_lu_batch(int nrep, const double *J, const double *mom, const double *lm0, int ref_mode);
double o39_quick_batch(int n, const double *rho, const double *T, const o39_consts *c);
int o39_lm_solve(double rho, const double u[3], double T, const double P[6], double lm[13],
                 double expo_last[39], int shifted, int ref_mode, int maxit);
double o39_alpha_estimate(const double f[39], const double diff[39], double T, int64_t step,
                          const o39_consts *c, int *status);
void o39_qbar(const double df[39], const double u[3], double T, double eps, int shifted,
              double qbar[39]);
void o39_little_qbar(const double dg[39], const double u[3], double T, double eps, int shifted,
                     double hq[39]);
int o39_collide(const double f[39], const double g[39], const double lm[13], double rho,
                const double u[3], double T, double alpha, int64_t step, const o39_consts *c,
                double fout[39], double gout[39]);
void o39_moments(const double f[39], const double g[39], int shifted, double *rho, double u[3],
                 double *T, double P[6], double cv);
int64_t o39_damp_periodic(int nx, int ny, int nz, double *fo, int shifted, int jacobi);
int64_t o39_damp_faces(int nx, int ny, int nz, const int32_t fc[6], double *fo, int shifted,
                       int jacobi);
void o39_init_fields(int nx, int ny, int nz, const double *rho, const double *ux, const double *uy,
                     const double *uz, const double *T, const o39_consts *c, double *f, double *g,
                     double *lm, int64_t *bad);
void o39_step_periodic(int nx, int ny, int nz, double *f, double *g, double *lm, double *alpha_out,
                       const o39_consts *c, int64_t step0, int nsteps, int ref_lu, int jacobi,
                       o39_stats *st);
void o39_step_ghost(int nx, int ny, int nz, int ghost, double *f, double *g, double *lm,
                    double *alpha_out, const o39_consts *c, int64_t step0, int nsteps, int ref_lu,
                    int jacobi, o39_stats *st);
int o39_source(int64_t a, int64_t b, int64_t cc, int k, int nx, int ny, int nz, const int32_t fc[6],
               int shifted, int64_t *src, int *dir);
/* as o39_source with inlet / outlet faces: face_row[f] is the table row of
 * face f (-2 the freestream pair, -1 pure outflow); returns 2 + row for a
 * table row */
int o39_source_bc(int64_t a, int64_t b, int64_t cc, int k, int nx, int ny, int nz,
                  const int32_t fc[6], const int32_t *face_row, int shifted, int64_t *src,
                  int *dir);
/* o_walls39.c: the shifted lattice's wall rule and a step with a link table */
/* wall-link gradients and rebuilt populations of the shifted-lattice wall
 * rule (o_walls39.c) */
void o39_wall_derivs(int i, const double ut[3], double su[7][3], const double sT[7],
                     const int scode[7], int shifted, double du[3][3], double dT[3]);
void o39_wall_grad(int i, double rho_t, const double ut[3], double Tt, double feq_i,
                   double du[3][3], const double dT[3], double eps, double b1, double b2,
                   const o39_consts *c, double *f_out, double *g_out);
int o39_wall_rebuild(int nx, int ny, int nz, const int32_t fc[6], const int32_t *state,
                     const double *f, const double *g, const double *lm, int64_t k,
                     const double fi[39], const o39_consts *c, int64_t step, int64_t lvl_step,
                     double fw[39], double gw[39], int *noconv, int *kerr);
/* plain-lattice wall rule (o_walls39.c): slip table, F/G of a direction,
 * the rebuilt populations of a node */
void o39p_set_slips(const double *slip, const int *start);
void o39p_wall_neq(int i, double T, const double ut[3], double du[3][3], const double dT[3],
                   double b1, double b2, double cv, double *F, double *G);
int o39p_wall_rebuild(int nx, int ny, int nz, const int32_t fc[6], const int32_t *state,
                      const double *f, const double *g, const double *lm, int64_t k,
                      const double fi[39], const o39_consts *c, int64_t step, double fw[39],
                      double gw[39], int *noconv, int *kerr);
void o39_step_walls(int nx, int ny, int nz, const int32_t fc[6], const int32_t *state,
                    const double *fs_f, const double *fs_g, const int32_t *face_row,
                    const double *bc_tab, double *f, double *g, double *lm,
                    double *alpha_out, const o39_consts *c, int64_t step0, int64_t lvl_step0,
                    int nsteps, int jacobi, o39_stats *st, int64_t *n_wall_links,
                    int64_t *n_wall_repair);
/* face_row / bc_tab: inlet and outlet faces (NULL when there are none): the
 * table holds rows of 39 f then 39 g values */
void o39_step_faces(int nx, int ny, int nz, const int32_t fc[6], const double *fs_f,
                    const double *fs_g, const int32_t *face_row, const double *bc_tab, double *f,
                    double *g, double *lm, double *alpha_out, const o39_consts *c, int64_t step0,
                    int nsteps, int ref_lu, int jacobi, o39_stats *st);

#ifdef __cplusplus
}
#endif
#endif
