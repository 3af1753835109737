/*
 * uclbs_hip.h -- C ABI of libuclbs_hip.so, the MI355X (gfx950) ELBM
 * collide-stream library that sits behind the UCLBS Fortran time loop.
 *
 * Boundary being replaced.  The reference advances one AMR level through six
 * Fortran phase calls in looping/timestepping/loop.f95:104-139,
 *     CALL streaming(level_step,fine,lvl)                 loop.f95:105
 *     CALL calculate_flow_variables(level_step,fine,lvl)  loop.f95:112
 *     CALL interlevel_interpolation(lvl,fine)             loop.f95:119
 *     CALL fine_to_coarse(lvl,fine)                       loop.f95:126
 *     CALL collision(level_step,fine,lvl)                 loop.f95:133
 *     CALL damp_it(level_step,fine,lvl)                   loop.f95:138
 * each of which sweeps the module-global AMReX MultiFabs of
 * amr/modules/amr_info_holder.f95:50-76 (fi, fout, gi, gout, state, rho,
 * u/v/w_vel, temp, alpha, the 13 Lagrange multipliers).  This library keeps
 * that level state resident in HBM and advances it with fused kernels; the
 * Fortran side reaches it through iso_c_binding (fortran/uclbs_hip_mod.f90).
 *
 * Conventions.
 *  - Every call returns 0 on success or a negative UCLBS_E_* code;
 *    uclbs_last_error() gives the message (the reference's error_out + STOP,
 *    general/background_calcs/error_out.f95:11-14, becomes a returned code).
 *  - Host arrays are FP64 in the reference layout: component-major,
 *    x fastest, i.e. an AMReX FAB fi(a,b,c,q) restricted to the box's valid
 *    nodes: f[q*n + (a-lo0) + nx*((b-lo1) + ny*(c-lo2))].
 *  - Velocity index q follows linkwise.f95:30-37 for D3Q27 (0-based).
 *  - The library owns one HIP stream per context; calls are synchronous
 *    unless their name ends in _async.
 *  - A context lives on the device that is current when it is created.
 */
#ifndef UCLBS_HIP_H
#define UCLBS_HIP_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

#define UCLBS_ABI_VERSION 1

enum {
  UCLBS_OK = 0,
  UCLBS_E_ARG = -1,      /* bad argument */
  UCLBS_E_HIP = -2,      /* HIP runtime error */
  UCLBS_E_STATE = -3,    /* call out of order (e.g. level not defined) */
  UCLBS_E_NOMEM = -4,    /* device allocation failed */
  UCLBS_E_NUMERIC = -5,  /* NaN or non-positive density detected */
  UCLBS_E_UNSUPPORTED = -6
};

enum { UCLBS_LATTICE_D3Q27 = 27, UCLBS_LATTICE_D3Q39 = 39 };
enum { UCLBS_PREC_FP32_SHIFTED = 0, UCLBS_PREC_FP64 = 1 };
enum {
  UCLBS_ALPHA_NEWTON = 0,   /* entropic Newton root-find from alpha = 2 (plain) */
  UCLBS_ALPHA_FIXED2 = 1,   /* alpha = 2 (LBGK), reference default turbulent=.false.
                               (calculate_flow_variables.f95:306-311) */
  UCLBS_ALPHA_LEARNED = 2   /* MLP initial guess, Newton, plain fallback */
};

/* Run configuration (the hot-path fields of the reference's &GLOBAL namelist,
 * general/namelist/namelist_global_read.f95:30-85). */
typedef struct uclbs_config {
  int32_t lattice;      /* UCLBS_LATTICE_* */
  int32_t precision;    /* UCLBS_PREC_* */
  int32_t alpha_mode;   /* UCLBS_ALPHA_* */
  int32_t alpha_maxit;  /* Newton cap (the reference has none,
                           calculate_lagrangian_mults.f95:341); 0 -> 30 */
  double alpha_tol;     /* |g(alpha)| <= tol * sum f x^2; 0 -> 1e-10 */
  double nu;            /* kinematic viscosity, lattice units: beta = 1/(6 nu + 1) */
  int32_t reserved[8];
} uclbs_config;

/* Counters accumulated on the device by every step. */
typedef struct uclbs_stats {
  int64_t node_updates;    /* nodes collided */
  int64_t newton_iters;    /* sum of Newton updates over nodes */
  int64_t n_trivial;       /* nodes at equilibrium (alpha = 2, no solve) */
  int64_t n_bound;         /* alpha clamped to the positivity bound */
  int64_t n_nonpos;        /* nodes with a non-positive population (alpha = 2) */
  int64_t n_noconv;        /* Newton hit the cap */
  int64_t n_fallback;      /* learned guess rejected, plain root-find used */
  int64_t n_logpath;       /* nodes whose alpha needed direct logarithms */
  int64_t n_ibb_links;     /* populations rebuilt by interpolated bounce-back */
  int64_t n_repair;        /* negative rebuilt populations replaced by |f|
                              (interpolative_bounceback_bc.f95:2969-2973) */
} uclbs_stats;

const char *uclbs_last_error(void);
int uclbs_abi_version(void);

/* Context: configuration and device stream.  Replaces the setup that
 * calcs_before_init/background_calcs perform (uclbs_main.f95:28-108). */
int uclbs_create(const uclbs_config *cfg, void **ctx);
int uclbs_destroy(void *ctx);

/* One rectangular nodal box per level: valid nodes lo..hi inclusive (AMReX
 * box convention, amr_processes.f95:383-466), periodic[d] != 0 wraps axis d.
 * Allocates two population buffers (A/B) in the storage precision. */
int uclbs_define_level(void *ctx, int lvl, const int32_t lo[3], const int32_t hi[3],
                       const int32_t periodic[3]);

/* Populations in and out (FP64 host, full f, reference layout). */
int uclbs_upload_f(void *ctx, int lvl, const double *f, int64_t n_nodes);
int uclbs_download_f(void *ctx, int lvl, double *f, int64_t n_nodes);
/* rho, u, v, w of the current populations (moments are conserved by the
 * collision, so these equal the fields calculate_flow_variables.f95:201-228
 * computes before the next collision). */
int uclbs_download_macro(void *ctx, int lvl, double *rho, double *u, double *v, double *w);
/* alpha of the last collision per node (requires uclbs_set_keep_alpha(1)). */
int uclbs_set_keep_alpha(void *ctx, int keep);
int uclbs_download_alpha(void *ctx, int lvl, double *alpha);

/* Initialisers (looping/initialization/init_fi.f95:41 fills fi/fout with an
 * equilibrium; the Taylor-Green field is the benchmark's own). */
int uclbs_init_tgv(void *ctx, int lvl, double u0, double noise_amp, uint64_t seed);

/* ---- boundaries and link tables ------------------------------------------ */
/* Face codes x-, y-, z-, x+, y+, z+ of a level's box, in the reference's
 * encoding of edge links (lbgrid/state_directions.f95:194-684): inlet j is
 * -100-j, outlet j is -200-j, -10 a bounce-back wall, -1000 freestream;
 * UCLBS_FACE_PERIODIC marks an axis declared periodic in uclbs_define_level.
 * Faces of non-periodic axes default to -10.
 * D3Q39 contexts (level 0) take -1000 (freestream_edge_bc.f95:18-19: the
 * missing population is the freestream equilibrium fs_fieq / fs_gieq),
 * UCLBS_FACE_OUTFLOW (an outlet of type 6, pure_outflow_bc.f95:17-20: the
 * node's own post-collision population of that direction), -10
 * (bounceback.f95:13-19: its opposite; the identity on the shifted lattice,
 * as the reference's opp) and UCLBS_FACE_INVISCID (-11, inviscid_edge.f95:
 * specular reflection about the half-way plane; not on the x faces of the
 * shifted lattice).  Where a population's upstream node lies across several
 * faces the first in the order x-, y-, z-, x+, y+, z+ decides, the order of
 * amr/amr_code/new_node_values/shifted_state_directions_lvl.f95:194-686. */
#define UCLBS_FACE_PERIODIC 2147483647
#define UCLBS_FACE_OUTFLOW (-6)
#define UCLBS_FACE_INVISCID (-11)
int uclbs_set_faces(void *ctx, int lvl, const int32_t face_code[6]);
/* Inlet / outlet / freestream states (rho, u, v, w) per boundary number; the
 * library tabulates their equilibria, as inlet_outlet_basis_calcs.f90:90-125
 * and freestream_basis_calcs.f90:35-194 tabulate inl_fieq / outl_fieq /
 * fs_fieq.  fs_state may be NULL. */
int uclbs_set_bc_states(void *ctx, int n_in, const double *in_states, int n_out,
                        const double *out_states, const double *fs_state);
/* Builds the level's link table on the device from a triangle soup in the
 * layout of geometry/modules/geom_data.f95:9-12 (tri[t*9 + vertex*3 + axis],
 * nrm[t*3 + axis]): node codes as lbgrid/fluid_or_solid.f95:86-442, wall-link
 * q codes as lbgrid/inter_bb_setup.f95:44-246 and linkwise.f95:45-57, face and
 * interior codes as lbgrid/state_directions.f95.  Node (a,b,c) of the box sits
 * at origin + (a,b,c)*dx (amr_processes.f95:1155-1168); on a refined D3Q27
 * level the box is the stored one (its ghost shell included) and dx the
 * level's spacing.  On a D3Q27 level of many boxes (refined levels) origin
 * is the position of level index (0,0,0): each box's stored nodes are built
 * at origin + (index)*dx, and a ghost node that shows a boundary or solid
 * node of another box is copied from it before each step.  The table stays
 * resident; *n_tie counts rays re-cast because of a tie.
 * On a D3Q39 context of the plain lattice the table has 39 components, with
 * the long links derived from the shorter ones as state_directions.f95:95-185
 * does (integer divisions and its slot slip for the (0,3,0) link included);
 * an outflow face gives its links the outside node's code 0.  Stepping
 * rebuilds the population opposite each wall link by the plain-lattice wall
 * rule (target_finder.f95:36-88, interpolative_bounceback_bc.f95:834-2838
 * with derivatives.f95:237-583: the Hermite projection its 38 cases expand
 * plus each case's departures from it).
 * On a D3Q39 context of the shifted lattice the table is that of
 * lbgrid/shifted_state_directions.f95: link i of node x is keyed on its
 * upstream node x - c_i (a wall code where that node is solid, the face code
 * where it lies across a face, the downstream node's code otherwise), and
 * stepping rebuilds the wall links by the shifted-lattice wall rule
 * (shifted_ibb.f95, target_finder.f95:87-127, derivatives.f95:816-1236). */
int uclbs_build_links(void *ctx, int lvl, const double origin[3], double dx, int64_t ntri,
                      const double *tri, const double *nrm, int64_t *n_solid, int64_t *n_tie);
/* Takes a reference-form table state[q*n + node] (int32, the `state` imultifab
 * of lbgrid/state_directions.f95:716 restricted to the box) instead. */
int uclbs_set_links(void *ctx, int lvl, const int32_t *state, int64_t n_nodes);
int uclbs_download_links(void *ctx, int lvl, int32_t *state, int64_t n_nodes);
/* fluid nodes with a boundary link, and solid/null nodes, of a level's table */
int uclbs_link_counts(void *ctx, int lvl, int64_t *n_boundary, int64_t *n_solid);
/* Uniform initial state: fluid nodes feq(rho, u), solid nodes the rest state
 * (looping/initialization/multilevel_initialization.f90:185-284). */
int uclbs_init_uniform(void *ctx, int lvl, double rho, double u, double v, double w);
/* Level-`level` icosphere (20*4^level triangles) of the given radius and
 * centre, rotated by a uniform random rotation drawn from Philox4x32-10 with
 * the seed; outward normals.  With tri or nrm NULL only *ntri is set.  A
 * generated stand-in for the STL import of geometry/geom_import.f95:48-133. */
int uclbs_geom_icosphere(int level, double radius, const double center[3], uint64_t seed,
                         double *tri, double *nrm, int64_t capacity, int64_t *ntri);
/* Symmetric wedge extruded along z: apex line at (apex_x, apex_y), faces at
 * +-half_angle_deg to the x axis, base at apex_x + length, spanning z0..z1.
 * Each side face is an n_len x n_span grid of triangle pairs whose interior
 * vertices are moved inside the face plane by Philox uniforms from the seed
 * (the surface is exact, the triangulation seeded); the end caps are fans over
 * the side faces' edge points, so the soup is watertight.  6*n_len*n_span +
 * 6*n_len triangles, outward normals; with tri or nrm NULL only *ntri is set.
 * A generated stand-in for the STL import of geometry/geom_import.f95:48-133
 * (the reference's supersonic wedge case). */
int uclbs_geom_wedge(double apex_x, double apex_y, double length, double half_angle_deg,
                     double z0, double z1, int n_len, int n_span, uint64_t seed, double *tri,
                     double *nrm, int64_t capacity, int64_t *ntri);

/* ---- D3Q39 two-population compressible model ----------------------------
 * A context created with lattice = UCLBS_LATTICE_D3Q39 advances f and g
 * (reference mffi/mfgi, amr_info_holder.f95:50-76) and the 13 Lagrange
 * multipliers per node (the warm start of calculate_lagrangian_mults).
 * Storage: FP32 deviations from the freestream equilibrium, or FP64.  Stats
 * for this lattice: newton_iters = multiplier Newton iterations, n_noconv =
 * Newton at its cap, n_trivial = alpha estimate 2 (calculate_alpha.f95:171-175,
 * 277-280), n_bound = estimate clamped (:281-289), n_nonpos = no real root
 * (the reference's error_out at :303-313), n_fallback = conductivity outside
 * the fitted bands (quick_calcs.f95:149-153, 207-212), n_repair =
 * populations repaired by damp_it.f95:64-92.  uclbs_step_level returns
 * UCLBS_E_NUMERIC when a step hit either error case. */
typedef struct uclbs_gas {
  double gama;          /* ratio of specific heats (namelist gama) */
  double temperature;   /* freestream temperature, K */
  double pressure;      /* freestream pressure, Pa */
  double mach;          /* freestream Mach number, primary flow along +x */
  double dx;            /* level-0 grid spacing, m (amrex_geom(0)%dx(1)) */
  int32_t startup_timesteps;  /* namelist startup_timesteps (reference default 25) */
  int32_t turbulent;    /* 1: alpha from calculate_alpha's estimate, 0: alpha = 2 */
  int32_t shifted;      /* -1: shift c_x by +1 when mach >= 1.2 (directions_and_weights.f95:477),
                           0 / 1: force */
  int32_t lm_maxit;     /* multiplier Newton cap per node (the reference has none); 0 -> 50 */
  int32_t reserved[4];
} uclbs_gas;
/* Derives the gas constants as basic_fluids_calcs.f95:17-86,
 * viscosity_and_timing.f95:29-73 and timing_calcs.f95:24-26 do.  Call before
 * uclbs_define_level on a D3Q39 context. */
int uclbs_set_gas(void *ctx, const uclbs_gas *gas);
/* Equilibrium state of the given fields (n_nodes FP64 each, layout as f):
 * multipliers by Newton from the Maxwellian's, f = f_eq, g = (2 c_v - 3) T f_eq. */
int uclbs_init_fields(void *ctx, int lvl, const double *rho, const double *u, const double *v,
                      const double *w, const double *temp);
/* f, g: [39][n] FP64 full values; lm: [13][n] multipliers (chi, zeta_x..z,
 * pi_xx, pi_yy, pi_zz, pi_xy, pi_xz, pi_yz, lambda_x..z), any may be NULL on download */
int uclbs_upload_state(void *ctx, int lvl, const double *f, const double *g, const double *lm,
                       int64_t n_nodes);
int uclbs_download_state(void *ctx, int lvl, double *f, double *g, double *lm, int64_t n_nodes);
/* ---- AMReX patch data on the device ------------------------------------------
 * The reference's level state lives in MultiFabs on the device
 * (amr_info_holder.f95:50-76: mffi, mfgi and the 13 multiplier fabs, nodal,
 * ghost width nghosts_mid).  uclbs_import_fab copies the valid region of such
 * a FAB -- FP64, component-major, every component the level's box grown by
 * `ghost` nodes on each side, x fastest, components comp_stride elements
 * apart -- into the level's state (populations converted to the storage form
 * on the device), with no host copy; uclbs_export_fab writes the state back
 * into the FAB's valid region (ghost cells untouched).  field: UCLBS_FIELD_F
 * (D3Q27 f or D3Q39 f), UCLBS_FIELD_G (D3Q39 g), UCLBS_FIELD_LM (D3Q39
 * multipliers, 13 components in the order of uclbs_upload_state).  The pointer
 * must be device memory of the context's device (checked); writes to it must
 * be complete before the call, and both calls return when the copy is done.
 * One-box levels that are not refined. */
enum { UCLBS_FIELD_F = 0, UCLBS_FIELD_G = 1, UCLBS_FIELD_LM = 2 };
int uclbs_import_fab(void *ctx, int lvl, int field, const double *fab, int ghost,
                     int64_t comp_stride);
int uclbs_export_fab(void *ctx, int lvl, int field, double *fab, int ghost, int64_t comp_stride);

/* ---- checkpoint / restart --------------------------------------------------
 * The reference's restart writer is commented out (restart_write.f95); these
 * give a versioned dump of the device state.  uclbs_checkpoint_write saves,
 * for every defined level, the current f (and g) buffer in storage form, the
 * D3Q39 multipliers and the step counter, with the level's set-up
 * fingerprint (extent, refined / many-box flags, face codes, D3Q39 gas
 * constants, a hash of the link table) and a checksum of the payload.
 * uclbs_checkpoint_read restores them into a context set up the same way
 * (same lattice and precision, same levels, faces, gas and links): it returns
 * UCLBS_E_STATE when the set-up differs and UCLBS_E_ARG for a file that is not
 * a checkpoint or is damaged; after an error the context's state is
 * unspecified.  A restored context continues bit for bit as the saved one. */
int uclbs_checkpoint_write(void *ctx, const char *path);
int uclbs_checkpoint_read(void *ctx, const char *path);
/* D3Q39 inlets and outlets with stored states (nml_inlet_outlet; face codes
 * -100-j for inlet j and -200-j for outlet j, j from 1).  in_states /
 * out_states hold (rho, u, v, w, T) per inlet / outlet, non-dimensional as the
 * reference derives them (inlet_outlet_base_values); types as the reference's
 * inlet_type / outlet_type: 1 rho-v, 2 pressure-temperature, 3 freestream,
 * 4 v-temperature, 7 (all four: the missing population is the equilibrium
 * of the stored state in its own direction, inlet_bc.f95:21-53 /
 * outlet_bc.f95:23-58; type 3 the freestream equilibrium, freestream_bc.f95),
 * and for outlets 6, pure outflow (pure_outflow_bc.f95).  The equilibria
 * come from the multiplier solve of each state from the reference's starting
 * guesses (inlet_outlet_basis_calcs.f90), with its two slips: g of every row
 * of a family uses the temperature of the last state solved after it, and an
 * outlet of type 2 reads the inlet table of its number (press_temp_bc.f95).
 * Needs uclbs_set_gas first. */
int uclbs_set_bc_states39(void *ctx, int n_in, const double *in_states, const int32_t *in_types,
                          int n_out, const double *out_states, const int32_t *out_types);
/* Flow variables of the current f, g per node, computed on the device as
 * calculate_flow_variables.f95:201-221 does: rho, u, v, w, temp and press =
 * rho (temp + |u|^2 / 2) (the reference's press, :221); any output may be
 * NULL.  n_nodes FP64 each, layout as f. */
int uclbs_download_macro39(void *ctx, int lvl, double *rho, double *u, double *v, double *w,
                           double *temp, double *press, int64_t n_nodes);
/* ---- refined levels ---------------------------------------------------------
 * uclbs_define_level with lvl >= 1 defines a box nested in level lvl-1 at
 * refinement ratio 2 (lo/hi in level lvl's own node indices; node G of level
 * lvl sits at coordinate G/2 of level lvl-1), stored with a one-node ghost
 * shell: its arrays (upload/download) hold (nx+2)(ny+2)(nz+2) nodes.  Lattice
 * viscosity doubles per level (acoustic scaling).  The ghost shell is filled
 * from the coarser level before each of the level's steps.
 * On a D3Q39 context the ghost layer is four nodes deep (links reach three
 * nodes, four along x on the shifted lattice; arrays (nx+8)(ny+8)(nz+8)): the
 * fill interpolates f, g and the coarse nodes' rho and T in space and time and
 * the multipliers in space, and replaces a negative f or g by the equilibrium
 * of the interpolated coarse state (interlevel_interpolation.f95:417-431);
 * uclbs_fine_to_coarse injects f, g and the multipliers; each level runs with
 * dt and epsilon halved per level (timing_calcs.f95:24-26). */
/* theta in [0, 1]: time between level lvl-1's previous and current states
 * (interlevel_interpolation.f95:114-493: amrex_fillcoarsepatch with
 * amrex_interp_node_bilinear; negative populations -> coarse equilibrium) */
int uclbs_interlevel(void *ctx, int lvl, double theta);
/* every node of level lvl from level lvl-1's current state (start-up / regrid) */
int uclbs_init_from_coarse(void *ctx, int lvl);
/* covered nodes of level lvl-1 take the coincident level-lvl populations
 * (fine_to_coarse.f95:25-65, amrex_average_down_nodes at ratio 2) */
int uclbs_fine_to_coarse(void *ctx, int lvl);

/* A level made of many disjoint boxes (the AMReX box array of a level,
 * amr_processes.f95:383-466): box b's valid nodes are lo[3b..3b+2] ..
 * hi[3b..3b+2] inclusive in the level's index space.  Each box is stored with
 * a one-node ghost shell, boxes one after another in the order given (box b
 * from node sum_{a<b} prod(hi_a - lo_a + 3), x fastest), and the whole level
 * advances in one launch of the batched streaming kernel.  Level 0: the boxes
 * must tile their bounding box, which is periodic on every axis.  Level l >= 1
 * nests in level l-1 (ratio 2): a ghost node that is no valid node of its own
 * level is interpolated from level l-1 by uclbs_interlevel (every such ghost
 * needs its coarse stencil inside level l-1's valid nodes), and
 * uclbs_fine_to_coarse injects the level's even nodes.  Host arrays of the
 * level (uclbs_upload_f, uclbs_download_f, uclbs_download_macro) are in this
 * stored order.  D3Q27.  Link tables: uclbs_build_links builds them box by
 * box (the level's origin is the position of index 0). */
int uclbs_define_level_boxes(void *ctx, int lvl, int nbox, const int32_t *lo, const int32_t *hi);

/* Advances level 0 by ncoarse steps and every finer level by two substeps per
 * step of its parent -- loop.f95:104-139 under the level schedule of
 * valid_levels.f95:21-59 -- with the phase order: step, ghosts of the finer
 * level at theta 0, its substep, ghosts at theta 1/2, its substep, injection.
 * use_graph != 0: the launches of two coarse steps are captured once as a
 * hipGraph and replayed (re-captured when the levels, kernel variants or
 * buffer parities change; an odd remainder runs eagerly).  On a D3Q39
 * context the start-up ramps read the finest level's step counter, so steps run eagerly
 * until every level is past its start-up. */
int uclbs_advance(void *ctx, int ncoarse, int use_graph);

/* Picks the tile (block) size of the batched streaming kernel of a level of
 * many boxes by timing 64, 128 and 256 nodes per block on the live state, then
 * restores the state; *tile_nodes receives the choice.  No counterpart in the
 * reference. */
int uclbs_autotune(void *ctx, int lvl, int *tile_nodes);

/* The level's step counter (reference timestep(lvl)), advanced by every step
 * of the level.  On a D3Q39 context the start-up ramps of beta, epsilon and
 * alpha of every level read the finest level's counter, the reference's
 * timestep(amr_max_lvl) (calculate_alpha.f95:185-187, collision.f95:277-278,
 * find_q_bar.f95:36, find_little_q_bar.f95:26). */
int uclbs_set_timestep(void *ctx, int lvl, int64_t step);

/* ---- learned alpha (D3Q27, FP32-shifted storage) --------------------------
 * No counterpart in the reference (whose closest precedent, the fitted
 * multiplier guesses of amr_processes.f95:1871-2039, is disabled at
 * calculate_lagrangian_mults.f95:729-741).  A 12-64-64-64-1 GELU network maps
 * 12 invariants of a node's non-equilibrium part to y = (alpha - 2) / m,
 * m = max|D_i / f_i|; with alpha_mode = UCLBS_ALPHA_LEARNED every step first
 * evaluates it on BF16 MFMA and the Newton root-find of H(f + alpha D) = H(f)
 * starts from alpha0 = 2 + m (y_mean + y_std y); the plain start alpha = 2
 * takes over when alpha0 lies outside the bracket or does not converge
 * (counted in uclbs_stats.n_fallback).  Parameters: W1 (12x64), b1, W2
 * (64x64), b2, W3 (64x64), b3, W4 (64), b4, FP32, 9,217 values. */
#define UCLBS_MLP_PARAMS 9217
/* uniform(+-sqrt(6/fan_in)) weights from Philox4x32-10 with the seed, zero biases */
int uclbs_mlp_init(void *ctx, uint64_t seed);
int uclbs_mlp_set_params(void *ctx, const float *params, float y_mean, float y_std);
int uclbs_mlp_get_params(void *ctx, float *params, float *y_mean, float *y_std);
/* loss 0.5 mean (net(X) - y)^2 and its gradient for a host batch X (B x 12), y (B) */
int uclbs_mlp_loss_grad(void *ctx, const float *X, const float *y, int B, float *grad,
                        double *loss);
/* network output for a host batch; use_mfma 1: BF16 MFMA path, 0: FP32 path */
int uclbs_mlp_forward(void *ctx, const float *X, int B, int use_mfma, float *yhat);
/* uclbs_mlp_forward with use_mfma = 1 for callers that cannot allocate device
 * memory (tight serving loops, harnesses that forbid hipMalloc).  No
 * counterpart in the reference.  It calls no hipMalloc / hipFree: the batch
 * is staged through a fixed area of the library (65,536 rows, 3.4 MB per
 * device, held until the process ends); B above 65,536 is refused with
 * UCLBS_E_ARG.  Every call on one device shares that area, so calls that
 * overlap on one device (two host threads, two contexts) race. */
int uclbs_mlp_forward_noalloc(void *ctx, const float *X, int B, float *yhat);
/* device sample buffer for (features, y) training pairs; capacity in
 * 1 .. 2^31 - 1 pairs (60 bytes of device memory per pair; collecting
 * adds 40 bytes per 256 nodes of the level) */
int uclbs_mlp_reserve(void *ctx, int64_t capacity);
/* logs every stride-th interior node of the current state, advances the level
 * one plain step (alpha_mode must be UCLBS_ALPHA_NEWTON) and stores the
 * resulting alpha as the target; *n_samples = pairs held.  Each call's pairs
 * are held in node order, so identical runs hold identical training sets
 * (pairs beyond the capacity are dropped) */
int uclbs_mlp_collect(void *ctx, int lvl, int stride, int64_t *n_samples);
int uclbs_mlp_download_samples(void *ctx, float *X, float *y, int64_t capacity, int64_t *n);
/* node index (in the level's FAB order) of each held pair, in the order of
 * uclbs_mlp_download_samples -- for checking the logged features */
int uclbs_mlp_download_sample_nodes(void *ctx, uint32_t *node, int64_t capacity, int64_t *n);
/* Adam (beta 0.9 / 0.999; peak rate lr after a linear warm-up over 2% of the
 * steps, cosine decay to 0.05 lr) on random batches of the held pairs
 * (targets normalised to zero mean, unit variance on first use) */
int uclbs_mlp_train(void *ctx, int64_t steps, int batch, double lr, uint64_t seed,
                    double *loss_first, double *loss_last);

/* Advance level lvl by nsteps fused stream+moments+alpha+collide steps:
 * the streaming / calculate_flow_variables / collision calls of
 * loop.f95:105,112,133 for one level, in that order. */
int uclbs_step_level(void *ctx, int lvl, int nsteps);
int uclbs_step_level_async(void *ctx, int lvl, int nsteps);
/* Same as uclbs_step_level, and returns the elapsed time of the launches as
 * measured by HIP events recorded on the library's stream. */
int uclbs_time_steps(void *ctx, int lvl, int nsteps, double *elapsed_ms);
/* As uclbs_time_steps; also records HIP events around each step's streaming
 * kernel (the dominant launch) and returns their summed duration. */
int uclbs_time_kernels(void *ctx, int lvl, int nsteps, double *step_ms, double *main_ms);
int uclbs_sync(void *ctx);

/* Measurement helper with no counterpart in the reference: bandwidth of a
 * plain copy of nbytes on the current device (read + write bytes / time), the
 * achievable ceiling reported next to the vendor HBM peak. */
int uclbs_copy_bandwidth(int64_t nbytes, int reps, double *gbps);

int uclbs_get_stats(void *ctx, uclbs_stats *out);
int uclbs_reset_stats(void *ctx);

/* ---- output: Plot3D files and microphones (output/plot3d_output.f95,
 * output/mic_setup_stuff/mic_setup.f95, mic_write.f95,
 * output/utilities/mic_initialize.f95) ---------------------------------- */

/* Physical placement of level 0: amrex_problo / amrex_probhi of the domain
 * (amr_processes.f95:1155-1168 get_real_coords).  Level l has spacing
 * (probhi - problo) / (cells of level 0 x 2^l) and its first domain node at
 * problo.  Without this call problo is 0 and level 0's spacing is 1. */
int uclbs_set_placement(void *ctx, const double problo[3], const double probhi[3]);

/* Header values of every solution block (plot3d_output.f95:256-258) */
typedef struct uclbs_plot3d_header {
  double mach, reynolds, time, gamma, t_ref, stag_temperature, cv, cp, gas_r;
  int32_t timestep;      /* names the files and ends each block header */
  int32_t max_level;     /* levels 0..max_level are written (clipped to the finest) */
  int64_t max_subrecord; /* bytes per Fortran subrecord; 0: gfortran's 2147483639 */
} uclbs_plot3d_header;

/* Replaces plot3d_output (plot3d_output.f95:76-305): writes
 * dir/uclbs_grid_<timestep, 9 digits>.x (Plot3D multi-block grid: block
 * count, extents, then per block x, y, z as REAL and IBLANK = 1 on fluid
 * nodes, -1 on code -666, 0 otherwise) and dir/uclbs_data_<timestep>.q
 * (extents with 6 variables, then per block the header record and rho, u,
 * v, w, T, alpha as FP64), Fortran unformatted sequential records.  One
 * block per box of each level, its valid region.  T is 1/3 on D3Q27;
 * alpha is the kept entropic alpha (uclbs_set_keep_alpha), else 2.
 * Asynchronous: the flow variables are computed and copied to pinned host
 * memory on the library's stream, a host thread writes the files; the
 * stream goes on stepping meanwhile.  uclbs_output_wait (or the next
 * uclbs_plot3d_write, or uclbs_destroy) waits for the files and reports a
 * failed write. */
int uclbs_plot3d_write(void *ctx, const char *dir, const uclbs_plot3d_header *h);
int uclbs_output_wait(void *ctx);

/* Replaces mic_setup + mic_initialize (mic_setup.f95:45-225,
 * mic_initialize.f95:17-70): n microphones at physical positions loc[n][3]
 * recording the fields of type (bits 32 pressure, 16 temperature,
 * 8 density, 4 u, 2 v, 1 w) on level min(level, finest).  Each gets the
 * eight corners of its cell and the reference's weights (distances with
 * + rdx on the far corners, divided by their sum over fluid corners; corners
 * at x index <= 0 weigh 0); files dir/microphoneNNNNN.txt get the header
 * line.  valid[i] (may be NULL): 0 invalid (outside [problo, probhi] or
 * no fluid corner), 1 sampled, 2 valid but outside the level's boxes (file
 * with the header only).  Set up after the levels and link tables. */
int uclbs_mic_setup(void *ctx, const char *dir, int n, const double *loc, const int32_t *type,
                    const int32_t *level, int32_t *valid);
/* Replaces mic_write (mic_write.f95:45-190) in two halves: _sample
 * enqueues the flow variables at every microphone's corners into a device
 * ring (no host synchronisation); _flush copies the held samples back and
 * appends one line per sample to each file (F12.5 values: p, T, rho, u, v, w
 * as selected).  A full ring flushes by itself. */
int uclbs_mic_sample(void *ctx);
int uclbs_mic_flush(void *ctx);
/* corners (level positions pos[8][3]), weights and stored nodes (-1: not a
 * fluid node) of microphone i (0-based) */
int uclbs_mic_weights(void *ctx, int i, int32_t pos[24], double w[8], int64_t node[8]);


/* ---- regrid -----------------------------------------------------------------
 * The dynamic half of the reference's AMR (regrid_question_mark.f95:38-184,
 * calling AMReX's regrid with the callbacks tag_nodes
 * (amr_processes.f95:549-664) and remake (amr_processes.f95:670-1080)):
 * tag a level by its GGM indicator, cluster the tags into boxes, remake the
 * next finer level on those boxes, then re-identify its nodes by building
 * its link table again (uclbs_build_links; the reference's
 * amr_code/new_node_values). */

/* Replaces remake (amr_processes.f95:670-1080): level lvl (>= 1, D3Q27) is
 * defined again as the nbox boxes lo/hi (as uclbs_define_level_boxes), every
 * node interpolated from level lvl-1 as uclbs_init_from_coarse does, then
 * every node at a position where the old level had a valid node takes the
 * old node's populations (both buffers).  The step counter carries over.
 * Finer levels become stale (stepping, filling or injecting them fails) until
 * each is regridded in turn, coarse to fine, or cleared.  When the old level
 * had a link table the new one must get its own (uclbs_build_links or
 * uclbs_set_links) before it steps.  On an error the old level stays. */
int uclbs_regrid_level(void *ctx, int lvl, int nbox, const int32_t *lo, const int32_t *hi);
/* nodes the last uclbs_regrid_level copied from the old level */
int uclbs_regrid_copied(void *ctx, int64_t *n);
/* Replaces clear_lvl (amr_processes.f95:478-545): frees level lvl (>= 1) and
 * every finer level. */
int uclbs_clear_level(void *ctx, int lvl);
/* Replaces ggm_calculator (GGM/ggm_calculator.f95:47-68: populate_q,
 * calculate_grad_mag, calculate_gradgrad_mag) on a D3Q27 level (one box, or a
 * refined level of many boxes, computed on a dense array over its stored
 * nodes' bounding box; a refined level's ghost shell is first filled from the
 * coarser level's current state): the
 * magnitude of the gradient of the gradient magnitude of a quantity
 * (1 density, 2 u, 3 v, 4 w, 7 velocity magnitude) from the current
 * populations, each derivative by derivative_windage_secondord
 * (derivatives.f95:644-700; scheme 0 central, 1 upwind, 2 downwind by the
 * node's velocity), stencil nodes usable where the node's link codes 1+a,
 * 4+a, 15+a, 18+a are >= 0 (every in-box node without a link table).  ggm
 * (may be NULL): [n] stored nodes; tags (may be NULL): 1 on the level's valid
 * nodes with ggm >= limit (calculate_gradgrad_mag.f95:67-71), 0 elsewhere;
 * ntagged (may be NULL): their count. */
int uclbs_tag_ggm(void *ctx, int lvl, int quantity, int scheme, double limit, double *ggm,
                  uint8_t *tags, int64_t *ntagged);
/* Boxes over tagged nodes (AMReX's cluster step between tag_nodes and
 * remake; host code, no device): tags [z][y][x] over lo..hi inclusive,
 * nonzero = tagged.  Tags grow by buffer nodes, coarsen by blocking, cluster
 * by Berger-Rigoutsos signatures until each box is at least efficiency
 * tagged, refine back, clip to lo..hi and split to at most max_size nodes per
 * axis (0: no limit; else a multiple of blocking).  Writes *nbox boxes
 * (disjoint, covering every grown tag) to box_lo/box_hi [nbox][3];
 * UCLBS_E_ARG with *nbox set when more than max_boxes are needed. */
int uclbs_cluster_tags(const uint8_t *tags, const int32_t lo[3], const int32_t hi[3],
                       double efficiency, int blocking, int buffer, int max_size, int max_boxes,
                       int32_t *box_lo, int32_t *box_hi, int *nbox);

#ifdef __cplusplus
}
#endif
#endif
