// D3Q39 kernels: fused pull-stream + moments + multiplier Newton + alpha
// estimate + collision for the two-population compressible model, the
// positivity repair, initialisation from macroscopic fields and layout
// conversion.
//
// Layout in HBM: f[q][node], g[q][node] (q = 0..38, node = x + nx(y + ny z)),
// FP32 deviations from the reference-state equilibrium (f - f_ref, g - g_ref)
// or FP64 full values; the 13 Lagrange multipliers lm[k][node] in FP64 (the
// warm start of the next step's Newton).  One thread per node, one wavefront
// per block: the node's pulled populations and the last Newton iteration's
// equilibrium sit in LDS, the 91-entry packed Jacobian in registers.  The
// step is FP64-bound: per node and Newton iteration 39 exponentials, a
// Jacobian of 39 x 91 FMAs (weights from a constant table, read as scalars)
// and a 13 x 13 L D L^T solve; alpha and the collision reuse the last
// iteration's equilibrium instead of recomputing 78 exponentials.
//
// Reference correspondence: streaming.f95:365-368 (pull form), then
// calculate_flow_variables.f95:201-311, calculate_alpha.f95:118-289,
// collision.f95:206-293 with find_q_bar.f95 / find_little_q_bar.f95 /
// quick_calcs.f95 / amr_processes.f95:1751-1814, and damp_it.f95:64-92.
#include "d3q39.h"
#include "ibb39_slips.h"
#include "kernels.h"

#include <type_traits>

namespace uclbs {

namespace {

__device__ __forceinline__ int wrapm(int s, int n) {
  s += (s < 0) ? n : 0;
  return s - ((s >= n) ? n : 0);
}

__device__ __forceinline__ void wave_stats39(unsigned long long* __restrict__ stats, bool active,
                                             int iters, bool noconv, int ast, bool kerr) {
  const unsigned long long act = __ballot(active);
  if (act == 0ull) return;
  unsigned long long it = 0;
#pragma unroll
  for (int bt = 0; bt < 8; ++bt)
    it += (unsigned long long)__popcll(__ballot(active && ((iters >> bt) & 1))) << bt;
  const unsigned long long nc = __popcll(__ballot(active && noconv));
  const unsigned long long nt = __popcll(__ballot(active && ast == 1));
  const unsigned long long nb = __popcll(__ballot(active && (ast == 2 || ast == 3)));
  const unsigned long long nn = __popcll(__ballot(active && ast < 0));
  const unsigned long long nk = __popcll(__ballot(active && kerr));
  if (__lane_id() == 0) {
    unsigned long long* s = stats + (size_t)(blockIdx.x & (kStatSlots - 1)) * kStatFields;
    atomicAdd(s + ST_NODES, (unsigned long long)__popcll(act));
    if (it) atomicAdd(s + ST_ITERS, it);
    if (nc) atomicAdd(s + ST_NOCONV, nc);
    if (nt) atomicAdd(s + ST_TRIVIAL, nt);
    if (nb) atomicAdd(s + ST_BOUND, nb);
    if (nn) atomicAdd(s + ST_NONPOS, nn);
    if (nk) atomicAdd(s + ST_FALLBACK, nk);
  }
}

// code of face f (lane-varying f: selects, not an indexed kernel-argument read)
__device__ __forceinline__ int face_code39(const Faces39& fc, int f) {
  int v = fc.code[0];
#pragma unroll
  for (int k = 1; k < 6; ++k) v = (f == k) ? fc.code[k] : v;
  return v;
}

__device__ __forceinline__ int face_tab39(const Faces39& fc, int f) {
  int v = fc.tab[0];
#pragma unroll
  for (int k = 1; k < 6; ++k) v = (f == k) ? fc.tab[k] : v;
  return v;
}

// Pull of population i into node (x, y, z) whose upstream node x - c_i may lie
// across a non-periodic face: streaming.f95:365-474 in pull form.  The first
// face crossed in the chain order -x, -y, -z, +x, +y, +z of
// shifted_state_directions_lvl.f95:194-686 decides: -1000 the freestream
// equilibrium (freestream_edge_bc.f95:18-19), -6 the node's own population
// (pure_outflow_bc.f95:17-20), -10 its opposite (bounceback.f95:13-19; opp is
// the identity on the shifted lattice), -11 the specular image about the
// half-way plane of every crossed inviscid face (inviscid_edge.f95:28-1243),
// after which a face still crossed decides as above; an inlet (-100-j) or an
// outlet (-200-j) the equilibrium of its stored state in direction i, or, for
// a pure-outflow outlet, the node's own population as -6.
template <bool kShifted, typename S>
__device__ __forceinline__ void pull_face39(int x, int y, int z, int i, const Box& b,
                                            const Faces39& fc, const Ref39& rf,
                                            const int (*smir)[Q39], const S* __restrict__ fA,
                                            const S* __restrict__ gA, S* fv, S* gv) {
  const int nb[3] = {b.nx, b.ny, b.nz};
  int s[3] = {x - cx39<kShifted>(i), y - CY39[i], z - CZ39[i]};
  int first = -1;
#pragma unroll
  for (int f = 5; f >= 0; --f) {
    const int d = f < 3 ? f : f - 3;
    const bool out = f < 3 ? s[d] < 0 : s[d] >= nb[d];
    if (out && fc.code[f] != kFacePeriodic) first = f;
  }
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    if (s[d] < 0 && fc.code[d] == kFacePeriodic) s[d] += nb[d];
    if (s[d] >= nb[d] && fc.code[d + 3] == kFacePeriodic) s[d] -= nb[d];
  }
  int dir = i;
  int code = first >= 0 ? face_code39(fc, first) : 0;
  if (code == -11) {
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      if (s[d] < 0 && fc.code[d] == -11) {
        s[d] = -1 - s[d];
        dir = smir[d][dir];
      } else if (s[d] >= nb[d] && fc.code[d + 3] == -11) {
        s[d] = 2 * nb[d] - 1 - s[d];
        dir = smir[d][dir];
      }
    }
    first = -1;
#pragma unroll
    for (int f = 5; f >= 0; --f) {
      const int d = f < 3 ? f : f - 3;
      if (f < 3 ? s[d] < 0 : s[d] >= nb[d]) first = f;
    }
    code = first >= 0 ? face_code39(fc, first) : 0;
  }
  const unsigned n = (unsigned)b.n;
  unsigned src;
  if (first < 0) {
    src = (unsigned)s[0] + (unsigned)nb[0] * ((unsigned)s[1] + (unsigned)nb[1] * (unsigned)s[2]);
  } else if (code == -1000) {
    *fv = (S)(fc.f[i] - rf.f[i]);
    *gv = (S)(fc.g[i] - rf.g[i]);
    return;
  } else if (code <= -101 && code >= -299 && face_tab39(fc, first) != -1) {
    // inlet / outlet with a stored state: its equilibrium in direction i
    // (inlet_bc.f95:21-53, outlet_bc.f95:23-58 via rho_v_bc / press_temp_bc /
    // v_temp_bc), or the freestream pair for type 3 (freestream_bc.f95)
    const int r = face_tab39(fc, first);
    if (r >= 0) {
      const double* row = fc.bc + (size_t)r * (2 * Q39);
      *fv = (S)(row[i] - rf.f[i]);
      *gv = (S)(row[Q39 + i] - rf.g[i]);
    } else {
      *fv = (S)(fc.f[i] - rf.f[i]);
      *gv = (S)(fc.g[i] - rf.g[i]);
    }
    return;
  } else {
    src = (unsigned)x + (unsigned)nb[0] * ((unsigned)y + (unsigned)nb[1] * (unsigned)z);
    dir = (code == -10 && !kShifted) ? OPP39[i] : i;
  }
  const S fr = fA[(size_t)dir * n + src], gr = gA[(size_t)dir * n + src];
  if (dir == i) {
    *fv = fr;
    *gv = gr;
  } else {  // FP32 storage holds deviations: re-base onto direction i's reference
    *fv = (S)(((double)fr + rf.f[dir]) - rf.f[i]);
    *gv = (S)(((double)gr + rf.g[dir]) - rf.g[i]);
  }
}

// calculate_alpha.f95:118-289 over the populations f (LDS, full value f_i =
// sf + fref) and the last Newton iteration's equilibrium (LDS, sd).  Returns
// alpha; *st = 0 (root), 1 (alpha = 2), 2 / 3 (clamped low / high), -1 (no
// real root: the reference stops).
template <bool kShifted, typename S>
__device__ __forceinline__ double alpha_estimate39(const S (*sf)[64], const double (*sd)[64], int t,
                                                   const Ref39& rf, double rho, double T,
                                                   long long step, const Gas39& gs, int* st) {
  const double alpha_hi = 2.125, alpha_start = 1.0e-6;
  double dim_mu, dim_temp;
  sutherland39(T, gs, &dim_mu, &dim_temp);
  const double kb = 1.38064852e-23, pi = 3.1415926535;
  const double mfp = dim_mu / gs.p_ref * sqrt(pi * kb * dim_temp / (2 * gs.molec_weight));
  const double mft = mfp / sqrt(3 * kb * dim_temp / gs.molec_weight);
  const double beta_mfp = gs.dt / (2 * mft + gs.dt);
  double cs1 = 0.0, cs2 = 0.0, cs3 = 0.0;
#pragma unroll 3
  for (int i = 0; i < Q39; ++i) {
    const double f = (double)sf[i][t] + rf.f[i];
    const double d = sd[i][t] - f;
    const double df = f < 0.0 ? fabs(f) : f;
    // x = d / df, df 2x^2/(2+x) and df x^2 over one division: e = df (2 + x)
    // is positive (the equilibrium is), r = 1 / (df e)
    const double e = 2.0 * df + d, r = 1.0 / (df * e), dd = d * d;
    const double x = d * e * r, q2 = dd * e * r;
    cs1 = cs1 + 2.0 * dd * df * r;
    cs2 = cs2 + q2;
    if (x <= 0.0 && x >= -1.0) cs3 = cs3 + q2 * x;
  }
  *st = 0;
  if (fabs(cs3) < 1.24e-16) {
    *st = 1;
    return 2.0;
  }
  double asm_ = (cs2 - sqrt(cs2 * cs2 - 8.0 * cs3 * cs1)) / (2.0 * cs3);
  if (asm_ > alpha_hi) asm_ = alpha_hi - 1.0e-9;
  if (step < gs.startup) asm_ = (asm_ - alpha_start) / (double)gs.startup * (double)step;
  double ft1 = 0, ft2 = 0, ft3 = 0, ft4 = 0, st1 = 0, st2 = 0, third = 0;
#pragma unroll 3
  for (int i = 0; i < Q39; ++i) {
    const double f = (double)sf[i][t] + rf.f[i];
    const double df = f < 0.0 ? fabs(f) : f;
    const double x1 = (sd[i][t] - f) / df, x2 = x1 * x1, x3 = x1 * x1 * x1;
    const double f2 = x2 * df, f3 = x3 * df, f4 = (x2 * x2) * df, f5 = x2 * x3 * df,
                 f6 = (x3 * x3) * df;
    if (x1 <= 0.0 && x1 >= -1.0) {
      ft1 += f3;
      ft2 += f4;
      ft3 += f5;
      ft4 += f6;
    }
    st1 += f2;
    // 2/D1 + 1/D2 + 2/D3 (x1 >= 0 only) and the third-order quotient share
    // one division: r = 1 / (B E), B = D1 D2 D3 or 1
    const double D1 = 4 + asm_ * x1, D2 = 4 + 2 * asm_ * x1, D3 = 4 + 3 * asm_ * x1;
    const bool pos = x1 >= 0.0;
    const double B = pos ? D1 * D2 * D3 : 1.0, E = 60 + 90 * x1 + 36 * x2 + 3 * x3;
    const double r = 1.0 / (B * E);
    if (pos) st2 += f3 * ((2.0 * D2 * D3 + D1 * D3 + 2.0 * D1 * D2) * E * r);
    third = third - (60 * f2 + 60 * f3 + 11 * f4) * B * r;
  }
  const double bm2 = beta_mfp * beta_mfp, b3 = bm2 * beta_mfp;
  const double a2h = alpha_hi * alpha_hi, a3h = a2h * alpha_hi;
  const double first = -bm2 * (ft1 / 6 - alpha_hi * beta_mfp * ft2 / 12 + a2h * bm2 * ft3 / 20 -
                               a3h * b3 * ft4 / 5);
  const double second = st1 / 2 - 2 * asm_ * bm2 * st2 / 15;
  if (fabs(first) < 1.0e-14) {
    *st = 1;
    return 2.0;
  }
  const double na = (-second + sqrt(second * second - 4 * first * third)) / (2 * first);
  if (!(na == na)) {
    *st = -1;
    return 2.0;
  }
  if (na < asm_) {
    *st = 2;
    return asm_;
  }
  if (na > alpha_hi) {
    *st = 3;
    return alpha_hi;
  }
  return na;
}

// kMode 0: every node of the box (near-face nodes skipped when the box has
// faces; with a link table, every node whose class is not 0); kMode 1: the
// listed near-face nodes, pulled through the face rules; kMode 2: the listed
// wall nodes, their wall links taken from the rebuilt values of
// k39_wall_rebuild and every other link pulled through the face rules.
template <bool kFp32, bool kShifted, bool kTurb, bool kKeepAlpha, int kMode>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(1, 2))) void k39_step(
    const void* __restrict__ fA_, const void* __restrict__ gA_, void* __restrict__ fB_,
    void* __restrict__ gB_, double* __restrict__ lmv, double* __restrict__ alpha, Box b,
    Gas39 gs, Ref39 rf, long long step, unsigned long long* __restrict__ stats,
    unsigned* __restrict__ neg_count, unsigned* __restrict__ neg_list, Faces39 fc, int has_faces,
    const unsigned* __restrict__ near_list, unsigned n_near, int ghost, Walls39 wl) {
  constexpr bool kNear = kMode != 0;
  using S = typename std::conditional<kFp32, float, double>::type;
  const S* __restrict__ fA = static_cast<const S*>(fA_);
  const S* __restrict__ gA = static_cast<const S*>(gA_);
  S* __restrict__ fB = static_cast<S*>(fB_);
  S* __restrict__ gB = static_cast<S*>(gB_);
  // staged populations (deviations from the reference equilibrium) in the
  // storage type; the wall nodes stage in FP64, since their rebuilt wall
  // populations sit far from the freestream and one FP32 rounding of those
  // would move the node's moments ~100x more than storage rounding does
  using SL = typename std::conditional<kMode == 2, double, S>::type;
  __shared__ SL sf[Q39][64];
  __shared__ SL sg[Q39][64];
  __shared__ double sd[Q39][64];
  __shared__ int smir[kNear ? 3 : 1][Q39];  // mirror directions for the inviscid faces
  const int t = threadIdx.x;
  if (kNear) {
    if (t < Q39) {
#pragma unroll
      for (int d = 0; d < 3; ++d) smir[d][t] = kMirror39.m[d][t];
    }
    __syncthreads();
  }
  const unsigned n = (unsigned)b.n;
  const unsigned idx = blockIdx.x * 64u + (unsigned)t;
  unsigned node;
  bool active;
  if (kMode == 1) {
    active = idx < n_near;
    node = active ? near_list[idx] : 0u;
  } else if (kMode == 2) {
    active = idx < wl.n;
    node = active ? wl.list[idx] : 0u;
  } else {
    node = idx;
    active = node < n;
    if (active && wl.cls) {
      active = wl.cls[node] == 0;
    } else if (active && (has_faces || ghost > 0)) {
      const int x = (int)(node % (unsigned)b.nx);
      const int y = (int)((node / (unsigned)b.nx) % (unsigned)b.ny);
      const int z = (int)(node / ((unsigned)b.nx * (unsigned)b.ny));
      if (has_faces) active = !near_face39(x, y, z, fc.code, b.nx, b.ny, b.nz, kShifted);
      // a refined level steps its valid nodes; its ghost layer (deeper than
      // the longest link) is filled from the coarser level
      if (ghost > 0)
        active = active && x >= ghost && y >= ghost && z >= ghost && x < b.nx - ghost &&
                 y < b.ny - ghost && z < b.nz - ghost;
    }
  }
  int iters = 0, ast = 0;
  bool noconv = false, kerr = false;
  if (active) {
    const int nx = b.nx, ny = b.ny, nz = b.nz;
    const int x = (int)(node % (unsigned)nx);
    const int y = (int)((node / (unsigned)nx) % (unsigned)ny);
    const int z = (int)(node / ((unsigned)nx * (unsigned)ny));
    if (kMode == 2) {
#pragma unroll 3
      for (int i = 0; i < Q39; ++i) {
        if (i > 0 && wl.state[(size_t)(kShifted ? i : OPP39[i]) * n + node] <= -1001) {
          sf[i][t] = wl.wf[(size_t)idx * Q39 + i] - rf.f[i];
          sg[i][t] = wl.wg[(size_t)idx * Q39 + i] - rf.g[i];
        } else {
          S fs, gsv;
          pull_face39<kShifted>(x, y, z, i, b, fc, rf, smir, fA, gA, &fs, &gsv);
          sf[i][t] = fs;
          sg[i][t] = gsv;
        }
      }
    } else if (kNear) {
#pragma unroll 3
      for (int i = 0; i < Q39; ++i) {
        S fs, gsv;
        pull_face39<kShifted>(x, y, z, i, b, fc, rf, smir, fA, gA, &fs, &gsv);
        sf[i][t] = fs;
        sg[i][t] = gsv;
      }
    } else {
#pragma unroll
      for (int i = 0; i < Q39; ++i) {
        const int cx = cx39<kShifted>(i), cy = CY39[i], cz = CZ39[i];
        const unsigned src = (unsigned)wrapm(x - cx, nx) +
                             (unsigned)nx * ((unsigned)wrapm(y - cy, ny) +
                                             (unsigned)ny * (unsigned)wrapm(z - cz, nz));
        sf[i][t] = fA[(size_t)i * n + src];
        sg[i][t] = gA[(size_t)i * n + src];
      }
    }
    double rho = 0.0, mx = 0.0, my = 0.0, mz = 0.0, en = 0.0;
#pragma unroll
    for (int i = 0; i < Q39; ++i) {
      const int cx = cx39<kShifted>(i), cy = CY39[i], cz = CZ39[i];
      const double f = (double)sf[i][t] + rf.f[i], g = (double)sg[i][t] + rf.g[i];
      const int cm = cx * cx + cy * cy + cz * cz;
      rho += f;
      if (cx) mx = fma((double)cx, f, mx);
      if (cy) my = fma((double)cy, f, my);
      if (cz) mz = fma((double)cz, f, mz);
      en += (double)cm * f + g;
    }
    double u[3] = {mx / rho, my / rho, mz / rho};
    const double vm = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
    const double T = (en - rho * vm) / (2 * gs.cv * rho);
    const double qf = rho * vm + 5 * rho * T;
    const double tg[kM39] = {rho, rho * u[0], rho * u[1], rho * u[2], rho * (T + u[0] * u[0]),
                             rho * (T + u[1] * u[1]), rho * (T + u[2] * u[2]), rho * u[0] * u[1],
                             rho * u[0] * u[2], rho * u[1] * u[2], qf * u[0], qf * u[1],
                             qf * u[2]};
    double lm[kM39], lmp[kM39];
#pragma unroll
    for (int k = 0; k < kM39; ++k) lm[k] = lmv[(size_t)k * n + node];
    // the last iteration's equilibrium stays in LDS for alpha and collision
    const int it = lm_newton39<kShifted>(rho, tg, lm, lmp, gs.maxit,
                                         [&](int i, double e) { sd[i][t] = e; });
    noconv = it < 0;
    iters = noconv ? gs.maxit : it;
#pragma unroll
    for (int k = 0; k < kM39; ++k) lmv[(size_t)k * n + node] = lm[k];

    double al = 2.0;
    if (kTurb) al = alpha_estimate39<kShifted>(sf, sd, t, rf, rho, T, step, gs, &ast);
    if (kKeepAlpha) alpha[node] = al;

    // collision: deviation from the updated multipliers' equilibrium,
    // exp(lm . m_i) = e_i exp(-(lm_prev - lm) . m_i) with e_i the last
    // iteration's value (D replaces it in LDS; the g deviation is rebuilt
    // from f + D where it is needed)
    double dl[kM39];
#pragma unroll
    for (int k = 0; k < kM39; ++k) dl[k] = lmp[k] - lm[k];
    double qxxx = 0, qyyy = 0, qzzz = 0, qxxy = 0, qxxz = 0, qxyy = 0, qxzz = 0, qyyz = 0,
           qyzz = 0, qxyz = 0, lq0 = 0, lq1 = 0, lq2 = 0;
    const double gT = gs.lead_g * T;
#pragma unroll
    for (int i = 0; i < Q39; ++i) {
      double sdl = dl[0];
#pragma unroll
      for (int k = 1; k < kM39; ++k) {
        const int m = mono39<kShifted>(i, k);
        if (m != 0) sdl += (double)m * dl[k];
      }
      const double feq = sd[i][t] * expm_small(sdl);
      const double f = (double)sf[i][t] + rf.f[i], g = (double)sg[i][t] + rf.g[i];
      const double d = feq - f;
      sd[i][t] = d;
      const double dgi = gT * feq - g;
      const double a = cx39<kShifted>(i) - u[0], bb = CY39[i] - u[1], e = CZ39[i] - u[2];
      qxxx += a * a * a * d;
      qyyy += bb * bb * bb * d;
      qzzz += e * e * e * d;
      qxxy += a * a * bb * d;
      qxxz += a * a * e * d;
      qxyy += bb * bb * a * d;
      qxzz += e * e * a * d;
      qyyz += bb * bb * e * d;
      qyzz += e * e * bb * d;
      qxyz += a * bb * e * d;
      lq0 += a * dgi;
      lq1 += bb * dgi;
      lq2 += e * dgi;
    }
    const double eps =
        step < gs.startup ? (((double)step + 1.0) / (double)gs.startup) * gs.eps : gs.eps;
    const double qs = eps / (6.0 * T * T * T), ql = eps / T;
    // contraction terms of qbar: sum_abc Q_abc c_a c_b c_c and sum_c (Q_aac) c_c
    const double sx = qxxx + qxyy + qxzz, sy = qxxy + qyyy + qyzz, sz = qxxz + qyyz + qzzz;
    double dmu, dte;
    const double mu = sutherland39(T, gs, &dmu, &dte);
    double kappa = 0.0;
    kerr = !therm_cond39(rho, T, gs, &kappa);
    double b1, b2;
    if (step < gs.startup) {
      b1 = gs.init_beta1 + gs.beta1_scale * (double)step;
      b2 = gs.init_beta2 + gs.beta2_scale * (double)step;
    } else {
      b1 = 1.0 / ((ldexp(1.0, gs.lvl + 1) * mu) / (rho * T) + 1.0);
      b2 = 1.0 / ((2.0 * kappa) / (rho * gs.cp * T) + 1.0);
    }
    const double ab1 = al * b1;
    bool neg = false;
#pragma unroll
    for (int i = 0; i < Q39; ++i) {
      const double cx = cx39<kShifted>(i), cy = CY39[i], cz = CZ39[i];
      const double cub = qxxx * cx * cx * cx + qyyy * cy * cy * cy + qzzz * cz * cz * cz +
                         3.0 * (qxxy * cx * cx * cy + qxxz * cx * cx * cz + qxyy * cx * cy * cy +
                                qxzz * cx * cz * cz + qyyz * cy * cy * cz + qyzz * cy * cz * cz) +
                         6.0 * qxyz * cx * cy * cz;
      const double qbar = qs * (cub - 3 * T * (sx * cx + sy * cy + sz * cz));
      const double hq = ql * (cx * lq0 + cy * lq1 + cz * lq2);
      const double f = (double)sf[i][t] + rf.f[i], g = (double)sg[i][t] + rf.g[i];
      const double d = sd[i][t];
      const double dgi = gT * (d + f) - g;
      const double fo = f + ab1 * (d - qbar) + b2 * qbar;
      const double go = g + ab1 * (dgi - hq) + b2 * hq;
      neg |= fo < 0.0;
      fB[(size_t)i * n + node] = (S)(fo - rf.f[i]);
      gB[(size_t)i * n + node] = (S)(go - rf.g[i]);
    }
    if (neg) {
      const unsigned slot = atomicAdd(neg_count, 1u);
      neg_list[slot] = node;
    }
  }
  wave_stats39(stats, active, iters, noconv, ast, kerr);
}

// ---- walls ---------------------------------------------------------------------
// rho, u, T of node k from the input buffers (the previous step's
// post-collision populations)
template <bool kShifted, typename S>
__device__ __forceinline__ void macro_at39(const S* __restrict__ fA, const S* __restrict__ gA,
                                           unsigned n, unsigned k, const Ref39& rf,
                                           const Gas39& gs, double* rho, double u[3], double* T) {
  double r = 0.0, mx = 0.0, my = 0.0, mz = 0.0, en = 0.0;
#pragma unroll 3
  for (int i = 0; i < Q39; ++i) {
    const int cx = cx39<kShifted>(i), cy = CY39[i], cz = CZ39[i];
    const double f = (double)fA[(size_t)i * n + k] + rf.f[i],
                 g = (double)gA[(size_t)i * n + k] + rf.g[i];
    r += f;
    mx += (double)cx * f;
    my += (double)cy * f;
    mz += (double)cz * f;
    en += (double)(cx * cx + cy * cy + cz * cz) * f + g;
  }
  u[0] = mx / r;
  u[1] = my / r;
  u[2] = mz / r;
  const double vm = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
  *rho = r;
  *T = (en - r * vm) / (2 * gs.cv * r);
}

// node index of (x, y, z) wrapped over the periodic faces; -1 beyond another
__device__ __forceinline__ long long pos39(int x, int y, int z, const Box& b, const Faces39& fc) {
  int s[3] = {x, y, z};
  const int nb[3] = {b.nx, b.ny, b.nz};
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    if (s[d] < 0) {
      if (fc.code[d] != kFacePeriodic) return -1;
      s[d] += nb[d];
    } else if (s[d] >= nb[d]) {
      if (fc.code[d + 3] != kFacePeriodic) return -1;
      s[d] -= nb[d];
    }
  }
  return (long long)s[0] + (long long)nb[0] * ((long long)s[1] + (long long)nb[1] * s[2]);
}

// The shifted lattice's wall rule for the listed wall nodes (streaming.f95:
// 478-1133, target_finder.f95:87-127, derivatives.f95:816-1236,
// shifted_ibb.f95:48-269), one thread per node: the target state (when a link
// of speed <= sqrt(3) is a wall link) -- u_t, T_t the q-weighted means over
// the wall links of the node's own u, T (or the downstream node's where the
// real opposite link is null), rho_t from this step's pulled populations
// with a wall link's population replaced by its real opposite's, the target
// multipliers by Newton from the node's (or, while the level's counter is
// <= 4, from fixed start-up guesses) -- else the node's own state; the
// velocity and temperature gradients from the 7-point stencil (wall-kind
// stencils from the reference's case table, WALLKIND39); then for every
// wall link f_i = feq_i(target) + eps k_f (P_comp + Q_comp) and
// g_i = (2 cv - 3) T_t feq_i + eps k_g (3 k1 + g_a + R_comp) with the
// reference's 27-entry tensors and term slips, a negative value replaced by
// its magnitude.  Fields of a node are the moments of its input-buffer
// populations; a stencil node beyond a non-periodic face counts as not fluid.
template <bool kFp32>
__global__ __launch_bounds__(64) void k39_wall_rebuild(const void* __restrict__ fA_,
                                                       const void* __restrict__ gA_,
                                                       const double* __restrict__ lmv, Box b,
                                                       Gas39 gs, Ref39 rf, Faces39 fc,
                                                       long long step, Walls39 wl,
                                                       unsigned long long* __restrict__ stats) {
  using S = typename std::conditional<kFp32, float, double>::type;
  const S* __restrict__ fA = static_cast<const S*>(fA_);
  const S* __restrict__ gA = static_cast<const S*>(gA_);
  __shared__ int smir[3][Q39];
  __shared__ double sfi[Q39][64];
  const int t = threadIdx.x;
  if (t < Q39) {
#pragma unroll
    for (int d = 0; d < 3; ++d) smir[d][t] = kMirror39.m[d][t];
  }
  __syncthreads();
  const unsigned j = blockIdx.x * 64u + (unsigned)t;
  if (j >= wl.n) return;
  const unsigned n = (unsigned)b.n;
  const unsigned node = wl.list[j];
  const int x = (int)(node % (unsigned)b.nx);
  const int y = (int)((node / (unsigned)b.nx) % (unsigned)b.ny);
  const int z = (int)(node / ((unsigned)b.nx * (unsigned)b.ny));
  const int32_t* __restrict__ st = wl.state;
  // this step's pulled populations (full values)
  for (int i = 0; i < Q39; ++i) {
    S fs, gsv;
    pull_face39<true>(x, y, z, i, b, fc, rf, smir, fA, gA, &fs, &gsv);
    sfi[i][t] = (double)fs + rf.f[i];
  }
  double rho0, u0[3], T0;
  macro_at39<true>(fA, gA, n, node, rf, gs, &rho0, u0, &T0);
  bool adjacent = false;
  for (int i = 1; i <= 14; ++i) adjacent = adjacent || st[(size_t)i * n + node] < -1000;
  double rho_t, ut[3], Tt, lmt[kM39];
  int noconv = 0;
  if (adjacent) {
    int count = 0;
    ut[0] = ut[1] = ut[2] = 0.0;
    Tt = 0.0;
    for (int i = 1; i < Q39; ++i) {
      const int32_t s = st[(size_t)i * n + node];
      if (s > -1001) continue;
      const double q = fabs((double)s + 1001.0) / 1.0e9;
      ++count;
      double rn = rho0, un[3] = {u0[0], u0[1], u0[2]}, Tn = T0;
      if (st[(size_t)OPP39[i] * n + node] == -666) {
        const long long d = pos39(x + cx39<true>(i), y + CY39[i], z + CZ39[i], b, fc);
        if (d >= 0) macro_at39<true>(fA, gA, n, (unsigned)d, rf, gs, &rn, un, &Tn);
      }
      for (int e = 0; e < 3; ++e) ut[e] += (q * un[e]) / (q + 1.0);
      Tt += (q * Tn + Tn) / (q + 1.0);
    }
    for (int e = 0; e < 3; ++e) ut[e] /= count;
    Tt /= count;
    rho_t = 0.0;
    for (int i = 0; i < Q39; ++i)
      rho_t += st[(size_t)i * n + node] < -1000 ? sfi[OPP39[i]][t] : sfi[i][t];
    if (wl.lvl_step <= 4) {
      // freestream_basis_calcs.f90:154-189
      const bool hi = wl.mach > 1.0;
      const double g0[kM39] = {hi ? -2.2 : -2.0, hi ? 0.4 : 0.0, 0.0, 0.0, hi ? -0.45 : -0.4,
                               hi ? -0.45 : -0.4, hi ? -0.45 : -0.4, 0.0, 0.0, 0.0,
                               hi ? 0.015 : 0.002, 0.0, 0.0};
      for (int k = 0; k < kM39; ++k) lmt[k] = g0[k];
    } else {
      for (int k = 0; k < kM39; ++k) lmt[k] = lmv[(size_t)k * n + node];
    }
    const double vm = ut[0] * ut[0] + ut[1] * ut[1] + ut[2] * ut[2];
    const double qf = rho_t * vm + 5 * rho_t * Tt;
    const double tg[kM39] = {rho_t, rho_t * ut[0], rho_t * ut[1], rho_t * ut[2],
                             rho_t * (Tt + ut[0] * ut[0]), rho_t * (Tt + ut[1] * ut[1]),
                             rho_t * (Tt + ut[2] * ut[2]), rho_t * ut[0] * ut[1],
                             rho_t * ut[0] * ut[2], rho_t * ut[1] * ut[2], qf * ut[0],
                             qf * ut[1], qf * ut[2]};
    double lmp[kM39];
    noconv = lm_newton39<true>(rho_t, tg, lmt, lmp, gs.maxit) < 0;
  } else {
    rho_t = rho0;
    for (int e = 0; e < 3; ++e) ut[e] = u0[e];
    Tt = T0;
    for (int k = 0; k < kM39; ++k) lmt[k] = lmv[(size_t)k * n + node];
  }
  // 7-point stencil: x, x + e_a, x - e_a
  double su[7][3], sT[7];
  int scode[7];
  for (int m = 0; m < 7; ++m) {
    const int a = m == 0 ? -1 : (m - 1) % 3, sg = m == 0 ? 0 : (m <= 3 ? 1 : -1);
    const long long d = m == 0 ? (long long)node
                               : pos39(x + (a == 0 ? sg : 0), y + (a == 1 ? sg : 0),
                                       z + (a == 2 ? sg : 0), b, fc);
    scode[m] = d >= 0 ? st[(size_t)d] : -1;
    if (scode[m] >= 0) {
      double rr, uu[3], TT;
      if (m == 0) {
        uu[0] = u0[0]; uu[1] = u0[1]; uu[2] = u0[2]; TT = T0;
      } else {
        macro_at39<true>(fA, gA, n, (unsigned)d, rf, gs, &rr, uu, &TT);
      }
      for (int e = 0; e < 3; ++e) su[m][e] = uu[e];
      sT[m] = TT;
    } else {
      su[m][0] = su[m][1] = su[m][2] = 0.0;
      sT[m] = 1.0 / gs.gama;
    }
  }
  const double eps =
      step < gs.startup ? (((double)step + 1.0) / (double)gs.startup) * gs.eps : gs.eps;
  double dmu, dte;
  const double mu = sutherland39(Tt, gs, &dmu, &dte);
  double kappa = 0.0;
  const bool kerr = !therm_cond39(rho_t, Tt, gs, &kappa);
  double b1, b2;
  if (step < gs.startup) {
    b1 = gs.init_beta1 + gs.beta1_scale * (double)step;
    b2 = gs.init_beta2 + gs.beta2_scale * (double)step;
  } else {
    b1 = 1.0 / ((ldexp(1.0, gs.lvl + 1) * mu) / (rho_t * Tt) + 1.0);
    b2 = 1.0 / ((2.0 * kappa) / (rho_t * gs.cp * Tt) + 1.0);
  }
  const double cv = gs.cv, T = Tt, u = ut[0], v = ut[1], w = ut[2];
  const double k_f = -(rho_t * T) / 2.0, k_g = -(rho_t * T * (2.0 * cv - 3.0)) / 2.0;
  unsigned long long nrep = 0, nlinks = 0;
  for (int i = 1; i < Q39; ++i) {
    if (st[(size_t)i * n + node] > -1001) continue;
    ++nlinks;
    const int X = cx39<true>(i), Y = CY39[i], Z = CZ39[i];
    const double link = sqrt((double)(X * X + Y * Y + Z * Z));
    double du[3][3], dT[3];
    for (int ax = 0; ax < 3; ++ax) {
      const int kind = WALLKIND39[i][ax];
      const double wd = fabs((double)scode[1 + ax] + 1001.0) / 1.0e9 * link;
      const bool up_ok = scode[4 + ax] >= 0, down_ok = scode[1 + ax] >= 0;
      for (int bb = 0; bb < 3; ++bb) {
        if (kind > 0) {
          du[ax][bb] = (0.0 - ut[bb]) / (1.0 + wd);
        } else if (kind < 0) {
          du[ax][bb] = (ut[bb] - 0.0) / (1.0 + wd);
        } else {
          const double rest = su[0][bb], up = su[4 + ax][bb], down = su[1 + ax][bb];
          du[ax][bb] = (up_ok && down_ok) ? (down - up) / 2.0
                       : up_ok            ? rest - up
                       : down_ok          ? down - rest
                                          : 0.0;
        }
      }
      if (kind != 0) {
        dT[ax] = 0.0;
      } else {
        const double rest = sT[0], up = sT[4 + ax], down = sT[1 + ax];
        dT[ax] = (up_ok && down_ok) ? (down - up) / 2.0
                 : up_ok            ? rest - up
                 : down_ok          ? down - rest
                                    : 0.0;
      }
    }
    const double dudx = du[0][0], dudy = du[1][0], dudz = du[2][0];
    const double dvdx = du[0][1], dvdy = du[1][1], dvdz = du[2][1];
    const double dwdx = du[0][2], dwdy = du[1][2], dwdz = du[2][2];
    const double dTdx = dT[0], dTdy = dT[1], dTdz = dT[2];
    double p[28];  // 1-based as in shifted_ibb.f95
    p[1] = 2 * dudx - dudx / cv; p[2] = 2 * dudx - dvdy / cv; p[3] = 2 * dudx - dwdz / cv;
    p[4] = p[5] = p[6] = dudy + dvdx;
    p[7] = p[8] = p[9] = dudz + dwdx;
    p[10] = p[11] = p[12] = dudy + dvdx;
    p[13] = 2 * dvdy - dudx / cv; p[14] = 2 * dvdy - dvdy / cv; p[15] = 2 * dvdy - dwdz / cv;
    p[16] = p[17] = p[18] = dvdz + dwdy;
    p[19] = p[20] = p[21] = dudz + dwdx;
    p[22] = p[23] = p[24] = dvdz + dwdy;
    p[25] = 2 * dwdz - dudx / cv; p[26] = 2 * dwdz - dvdy / cv; p[27] = 2 * dwdz - dwdz / cv;
    for (int m = 1; m <= 27; ++m) p[m] = p[m] / b1;
    const double tx = dTdx / b2, ty = dTdy / b2, tz = dTdz / b2;
    double qq[28];
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
    const double Xd = X, Yd = Y, Zd = Z;
    const double sxx = Xd * Xd - T, syy = Yd * Yd - T, szz = Zd * Zd - T;
    const double pc[28] = {0, sxx, sxx, sxx, Xd * Yd, Xd * Yd, Xd * Yd, Xd * Zd, Xd * Zd, Xd * Zd,
                           Yd * Xd, Yd * Xd, Yd * Xd, syy, syy, syy, Yd * Zd, Yd * Zd, Yd * Zd,
                           Zd * Xd, Zd * Xd, Zd * Xd, Zd * Yd, Zd * Yd, Zd * Yd, szz, szz, szz};
    double psum = 0.0;
    for (int m = 1; m <= 27; ++m) psum += p[m] * pc[m];
    psum += p[25] * szz;  // P_zzx counted twice (shifted_ibb.f95:170)
    const double p_comp = psum / (2 * T * T);
    const double qc[28] = {0,
                           Xd * Xd * Xd - 3 * Xd * T, Xd * Xd * Yd - 3 * Yd * T,
                           Xd * Xd * Zd - 3 * Zd * T, Xd * Xd * Yd, Xd * Yd * Yd, Xd * Yd * Zd,
                           Xd * Xd * Zd, Xd * Yd * Zd, Xd * Zd * Zd, Xd * Xd * Yd, Yd * Yd * Xd,
                           Xd * Yd * Zd, Yd * Yd * Xd - 3 * Xd * T, Yd * Yd * Yd - 3 * Yd * T,
                           Yd * Yd * Zd - 3 * Zd * T, Xd * Yd * Zd, Yd * Yd * Zd, Zd * Zd * Yd,
                           Xd * Xd * Zd, Xd * Yd * Zd, Zd * Zd * Xd, Xd * Yd * Zd, Zd * Yd * Yd,
                           Zd * Zd * Yd, Zd * Zd * Xd - 3 * Xd * T, Zd * Zd * Yd - 3 * Yd * T,
                           Zd * Zd * Zd - 3 * Zd * T};
    double qsum = 0.0;
    for (int m = 1; m <= 27; ++m) qsum += qq[m] * qc[m];
    const double q_comp = qsum / (6 * T * T * T);
    const double feq = expo39<true>(lmt, rho_t, i);
    double fo = feq + eps * k_f * (p_comp + q_comp);
    const double k_one = 3 * (dTdx + dTdy + dTdz) / (b1 * cv);
    const double g_alpha = Xd / T * (9 * dTdx / b2 + u * k_one) + Yd / T * (9 * dTdy / b2 + v * k_one) +
                           Zd / T * (9 * dTdz / b2 + w * k_one);
    double r[28];
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
    r[22] = r[13]; r[23] = r[14]; r[24] = r[15];  // the zy entries repeat yy (shifted_ibb.f95:231-233)
    r[25] = 2 * T * dwdz + 2 * w * dTdz - dudx / cv * (2 * T + w * w);
    r[26] = 2 * T * dwdz + 2 * w * dTdz - dvdy / cv * (2 * T + w * w);
    r[27] = 2 * T * dwdz + 2 * w * dTdz - dwdz / cv * (2 * T + w * w);
    double rsum = 0.0;
    for (int m = 1; m <= 27; ++m) rsum += r[m] * pc[m];
    const double gi_neq = k_g * (3 * k_one + g_alpha + rsum / (2 * b1 * T));
    double go = gs.lead_g * T * feq + eps * gi_neq;
    if (fo < 0.0) { fo = fabs(fo); ++nrep; }
    if (go < 0.0) { go = fabs(go); ++nrep; }
    wl.wf[(size_t)j * Q39 + i] = fo;
    wl.wg[(size_t)j * Q39 + i] = go;
  }
  unsigned long long* sl = stats + (size_t)(blockIdx.x & (kStatSlots - 1)) * kStatFields;
  atomicAdd(sl + ST_IBB, nlinks);
  if (nrep) atomicAdd(sl + ST_REPAIR, nrep);
  if (noconv) atomicAdd(sl + ST_NOCONV, 1ull);
  if (kerr) atomicAdd(sl + ST_FALLBACK, 1ull);
}

// x^e for a small integer e
__device__ __forceinline__ double ipow39(double x, int e) {
  double r = 1.0;
  for (int k = 0; k < (e < 0 ? -e : e); ++k) r *= x;
  return e < 0 ? 1.0 / r : r;
}

// Non-equilibrium parts F_i, G_i of the plain lattice's rebuilt population i
// (1..38): the Hermite projection the reference's hand-expanded cases of
// interpolative_bounceback_bc.f95:834-2838 carry out,
//   F = H2(c):P / (2 b1 T^2) + H3(c):Q / (6 T^3),
//   G = 9 div / (b1 cv) + c.(9 grad T / b2 + 3 u div / (b1 cv)) / T
//       + H2(c):R / (2 b1 T^2),
// P_ab = 3 (d_a u_b + d_b u_a) - delta_ab div / cv, Q_abe the sum over the
// three pair/single splits (p q | s) of delta_pq d_s T / b2 + u_s Pi_pq^s / b1
// with Pi_pp^s = 2 d_p u_p - d_s u_s / cv and Pi_pq = d_p u_q + d_q u_p,
// R_aa = 6 T d_a u_a + 6 u_a d_a T - (2 T + u_a^2) div / cv and
// R_ab = 3 (T (d_a u_b + d_b u_a) + u_a d_b T + u_b d_a T) - u_a u_b div / cv;
// plus the terms by which the cases depart from it (kIbbSlip).
// du[a][b] = d u_b / d x_a.
__device__ void ibb_neq_plain39(int i, double T, const double ut[3], const double du[3][3],
                                const double dT[3], double b1, double b2, double cv, double* F,
                                double* G) {
  const int c[3] = {cx39<false>(i), CY39[i], CZ39[i]};
  const double div = du[0][0] + du[1][1] + du[2][2];
  double h2p = 0.0, h2r = 0.0, h3q = 0.0;
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b) {
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
      for (int e = 0; e < 3; ++e) {
        const double h3 = (double)(c[a] * c[b] * c[e]) -
                          T * (double)(c[a] * (b == e) + c[b] * (a == e) + c[e] * (a == b));
        if (h3 == 0.0) continue;
        const int sp[3][3] = {{a, b, e}, {a, e, b}, {b, e, a}};
        double q = 0.0;
        for (int m = 0; m < 3; ++m) {
          const int p = sp[m][0], r = sp[m][1], sg = sp[m][2];
          const double pi = p == r ? 2.0 * du[p][p] - du[sg][sg] / cv : du[p][r] + du[r][p];
          q += (p == r ? dT[sg] / b2 : 0.0) + ut[sg] * pi / b1;
        }
        h3q += h3 * q;
      }
    }
  double f = h2p / (2.0 * b1 * T * T) + h3q / (6.0 * T * T * T);
  double g = 9.0 * div / (b1 * cv) + h2r / (2.0 * b1 * T * T);
  for (int a = 0; a < 3; ++a) g += (double)c[a] * (9.0 * dT[a] / b2 + 3.0 * ut[a] * div / (b1 * cv)) / T;
  const double D[12] = {du[0][0], du[1][0], du[2][0], du[0][1], du[1][1], du[2][1],
                        du[0][2], du[1][2], du[2][2], dT[0], dT[1], dT[2]};
  for (int part = 0; part < 2; ++part) {
    double add = 0.0;
    for (int m = kIbbSlipStart[38 * part + i - 1]; m < kIbbSlipStart[38 * part + i]; ++m) {
      const IbbSlip& t = kIbbSlip[m];
      add += t.coef * D[t.d] / (t.beta == 1 ? b1 : b2) * ipow39(ut[0], t.eu) * ipow39(ut[1], t.ev) *
             ipow39(ut[2], t.ew) * ipow39(T, t.eT) * ipow39(cv, t.ecv);
    }
    if (part == 0) f += add;
    else g += add;
  }
  *F = f;
  *G = g;
}

// The plain lattice's wall rule for the listed wall nodes (streaming.f95:
// 478-1133 with the plain branch, target_finder.f95:36-88,
// interpolative_bounceback_bc.f95:834-2838, derivatives.f95:237-583), one
// thread per node.  Link j (x -> x + c_j) coded <= -1001 makes the population
// i = opp(j), which would be pulled from the solid node x + c_j, a rebuilt
// one.  Target (when a link of speed <= sqrt(3) is a wall link): u_t, T_t
// the means over the wall links of q u_n / (q + 1) and (q T_n + T_n) /
// (q + 1), n = x + c_opp(j) when that link's code is >= 0 (the node itself
// beyond a face), else q u(x) / (q + 1) and (q T(x) + T_inf) / (q + 1);
// rho_t from this step's pulled populations with a missing one replaced by
// its opposite; multipliers by Newton from the node's; otherwise the node's
// own state.  Gradients: along an axis where c_i is positive
// (u_t - 0) / (1 + w), negative (0 - u_t) / (1 + w), w = |s_i + 1001| 1e-9
// |c_i| from the node's own code of link i, temperature 0; else the local
// difference over the neighbours whose link code is >= 0.  Then
// f_i = feq_i(target) + eps k_f F_i, g_i = (2 cv - 3) T_t feq_i + eps k_g G_i
// (ibb_neq_plain39); no magnitude repair (the repair pass handles negatives).
template <bool kFp32>
__global__ __launch_bounds__(64) void k39_wall_rebuild_plain(const void* __restrict__ fA_,
                                                             const void* __restrict__ gA_,
                                                             const double* __restrict__ lmv,
                                                             Box b, Gas39 gs, Ref39 rf,
                                                             Faces39 fc, long long step,
                                                             Walls39 wl,
                                                             unsigned long long* __restrict__ stats) {
  using S = typename std::conditional<kFp32, float, double>::type;
  const S* __restrict__ fA = static_cast<const S*>(fA_);
  const S* __restrict__ gA = static_cast<const S*>(gA_);
  __shared__ int smir[3][Q39];
  __shared__ double sfi[Q39][64];
  const int t = threadIdx.x;
  if (t < Q39) {
#pragma unroll
    for (int d = 0; d < 3; ++d) smir[d][t] = kMirror39.m[d][t];
  }
  __syncthreads();
  const unsigned j = blockIdx.x * 64u + (unsigned)t;
  if (j >= wl.n) return;
  const unsigned n = (unsigned)b.n;
  const unsigned node = wl.list[j];
  const int x = (int)(node % (unsigned)b.nx);
  const int y = (int)((node / (unsigned)b.nx) % (unsigned)b.ny);
  const int z = (int)(node / ((unsigned)b.nx * (unsigned)b.ny));
  const int32_t* __restrict__ st = wl.state;
  for (int i = 0; i < Q39; ++i) {
    S fs, gsv;
    pull_face39<false>(x, y, z, i, b, fc, rf, smir, fA, gA, &fs, &gsv);
    sfi[i][t] = (double)fs + rf.f[i];
  }
  double rho0, u0[3], T0;
  macro_at39<false>(fA, gA, n, node, rf, gs, &rho0, u0, &T0);
  const double t_inf = 1.0 / gs.gama;
  bool adjacent = false;
  for (int i = 1; i <= 14; ++i) adjacent = adjacent || st[(size_t)i * n + node] < -1000;
  double rho_t, ut[3], Tt, lmt[kM39];
  int noconv = 0;
  for (int k = 0; k < kM39; ++k) lmt[k] = lmv[(size_t)k * n + node];
  if (adjacent) {
    int count = 0;
    ut[0] = ut[1] = ut[2] = 0.0;
    Tt = 0.0;
    for (int i = 1; i < Q39; ++i) {
      const int32_t s = st[(size_t)i * n + node];
      if (s > -1001) continue;
      const double q = fabs((double)s + 1001.0) / 1.0e9;
      ++count;
      const int o = OPP39[i];
      if (st[(size_t)o * n + node] >= 0) {
        double rn = rho0, un[3] = {u0[0], u0[1], u0[2]}, Tn = T0;
        const long long d = pos39(x + cx39<false>(o), y + CY39[o], z + CZ39[o], b, fc);
        if (d >= 0) macro_at39<false>(fA, gA, n, (unsigned)d, rf, gs, &rn, un, &Tn);
        for (int e = 0; e < 3; ++e) ut[e] += (q * un[e]) / (q + 1.0);
        Tt += (q * Tn + Tn) / (q + 1.0);
      } else {
        for (int e = 0; e < 3; ++e) ut[e] += (q * u0[e]) / (q + 1.0);
        Tt += (q * T0 + t_inf) / (q + 1.0);
      }
    }
    for (int e = 0; e < 3; ++e) ut[e] /= count;
    Tt /= count;
    rho_t = 0.0;
    for (int i = 0; i < Q39; ++i)
      rho_t += st[(size_t)OPP39[i] * n + node] < -1000 ? sfi[OPP39[i]][t] : sfi[i][t];
    const double vm = ut[0] * ut[0] + ut[1] * ut[1] + ut[2] * ut[2];
    const double qf = rho_t * vm + 5 * rho_t * Tt;
    const double tg[kM39] = {rho_t, rho_t * ut[0], rho_t * ut[1], rho_t * ut[2],
                             rho_t * (Tt + ut[0] * ut[0]), rho_t * (Tt + ut[1] * ut[1]),
                             rho_t * (Tt + ut[2] * ut[2]), rho_t * ut[0] * ut[1],
                             rho_t * ut[0] * ut[2], rho_t * ut[1] * ut[2], qf * ut[0],
                             qf * ut[1], qf * ut[2]};
    double lmp[kM39];
    noconv = lm_newton39<false>(rho_t, tg, lmt, lmp, gs.maxit) < 0;
  } else {
    rho_t = rho0;
    for (int e = 0; e < 3; ++e) ut[e] = u0[e];
    Tt = T0;
  }
  // stencil values: x, x + e_a, x - e_a (a fluid node's own; 0 and T_inf else)
  double su[7][3], sT[7];
  for (int m = 0; m < 7; ++m) {
    const int a = m == 0 ? -1 : (m - 1) % 3, sg = m == 0 ? 0 : (m <= 3 ? 1 : -1);
    const long long d = m == 0 ? (long long)node
                               : pos39(x + (a == 0 ? sg : 0), y + (a == 1 ? sg : 0),
                                       z + (a == 2 ? sg : 0), b, fc);
    if (d >= 0 && st[(size_t)d] >= 0) {
      double rr, uu[3], TT;
      if (m == 0) {
        uu[0] = u0[0]; uu[1] = u0[1]; uu[2] = u0[2]; TT = T0;
      } else {
        macro_at39<false>(fA, gA, n, (unsigned)d, rf, gs, &rr, uu, &TT);
      }
      for (int e = 0; e < 3; ++e) su[m][e] = uu[e];
      sT[m] = TT;
    } else {
      su[m][0] = su[m][1] = su[m][2] = 0.0;
      sT[m] = t_inf;
    }
  }
  const double eps =
      step < gs.startup ? (((double)step + 1.0) / (double)gs.startup) * gs.eps : gs.eps;
  double dmu, dte;
  const double mu = sutherland39(Tt, gs, &dmu, &dte);
  double kappa = 0.0;
  const bool kerr = !therm_cond39(rho_t, Tt, gs, &kappa);
  double b1, b2;
  if (step < gs.startup) {
    b1 = gs.init_beta1 + gs.beta1_scale * (double)step;
    b2 = gs.init_beta2 + gs.beta2_scale * (double)step;
  } else {
    b1 = 1.0 / ((ldexp(1.0, gs.lvl + 1) * mu) / (rho_t * Tt) + 1.0);
    b2 = 1.0 / ((2.0 * kappa) / (rho_t * gs.cp * Tt) + 1.0);
  }
  const double k_f = -(rho_t * Tt) / 2.0, k_g = -(rho_t * Tt * (2.0 * gs.cv - 3.0)) / 2.0;
  unsigned long long nlinks = 0;
  for (int i = 1; i < Q39; ++i) {
    if (st[(size_t)OPP39[i] * n + node] > -1001) continue;
    ++nlinks;
    const int ci[3] = {cx39<false>(i), CY39[i], CZ39[i]};
    const double link = sqrt((double)(ci[0] * ci[0] + ci[1] * ci[1] + ci[2] * ci[2]));
    const double wd = fabs((double)st[(size_t)i * n + node] + 1001.0) / 1.0e9 * link;
    double du[3][3], dT[3];
    for (int ax = 0; ax < 3; ++ax) {
      const bool up_ok = st[(size_t)(4 + ax) * n + node] >= 0;
      const bool down_ok = st[(size_t)(1 + ax) * n + node] >= 0;
      for (int e = 0; e < 3; ++e) {
        if (ci[ax] > 0) {
          du[ax][e] = (ut[e] - 0.0) / (1.0 + wd);
        } else if (ci[ax] < 0) {
          du[ax][e] = (0.0 - ut[e]) / (1.0 + wd);
        } else {
          const double rest = su[0][e], up = su[4 + ax][e], down = su[1 + ax][e];
          du[ax][e] = (up_ok && down_ok) ? (down - up) / 2.0
                      : up_ok            ? rest - up
                      : down_ok          ? down - rest
                                         : 0.0;
        }
      }
      if (ci[ax] != 0) {
        dT[ax] = 0.0;
      } else {
        const double rest = sT[0], up = sT[4 + ax], down = sT[1 + ax];
        dT[ax] = (up_ok && down_ok) ? (down - up) / 2.0
                 : up_ok            ? rest - up
                 : down_ok          ? down - rest
                                    : 0.0;
      }
    }
    double F, G;
    ibb_neq_plain39(i, Tt, ut, du, dT, b1, b2, gs.cv, &F, &G);
    const double feq = expo39<false>(lmt, rho_t, i);
    wl.wf[(size_t)j * Q39 + i] = feq + eps * k_f * F;
    wl.wg[(size_t)j * Q39 + i] = gs.lead_g * Tt * feq + eps * k_g * G;
  }
  unsigned long long* sl = stats + (size_t)(blockIdx.x & (kStatSlots - 1)) * kStatFields;
  atomicAdd(sl + ST_IBB, nlinks);
  if (noconv) atomicAdd(sl + ST_NOCONV, 1ull);
  if (kerr) atomicAdd(sl + ST_FALLBACK, 1ull);
}

// node (x, y, z) wrapped over the periodic axes of b; -1 across another face
// or inside a refined level's ghost layer (which holds no post-collision values)
__device__ __forceinline__ long long damp_nb(int x, int y, int z, const Box& b, int ghost) {
  int s[3] = {x, y, z};
  const int nb[3] = {b.nx, b.ny, b.nz};
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    if (s[d] < 0 || s[d] >= nb[d]) {
      if (!b.per[d]) return -1;
      s[d] = wrapm(s[d], nb[d]);
    }
    if (ghost > 0 && (s[d] < ghost || s[d] >= nb[d] - ghost)) return -1;
  }
  return (long long)s[0] + (long long)nb[0] * ((long long)s[1] + (long long)nb[1] * s[2]);
}

// damp_it.f95:64-92 for the listed nodes: each negative post-collision
// population takes the mean of its positive neighbours along +-c_i (or the
// mean of one with |f|, or |f|), all read before any repair is written;
// the new values go to rep[j][q] (NaN where nothing changes).
template <bool kFp32, bool kShifted>
__global__ __launch_bounds__(256) void k39_damp_compute(const void* __restrict__ f_, Box b,
                                                        Ref39 rf, const unsigned* __restrict__ count,
                                                        const unsigned* __restrict__ list,
                                                        double* __restrict__ rep,
                                                        unsigned long long* __restrict__ stats,
                                                        int ghost, const int32_t* __restrict__ state) {
  using S = typename std::conditional<kFp32, float, double>::type;
  const S* __restrict__ f = static_cast<const S*>(f_);
  const unsigned cnt = *count;
  const unsigned n = (unsigned)b.n;
  unsigned long long nrep = 0;
  for (unsigned j = blockIdx.x * blockDim.x + threadIdx.x; j < cnt; j += gridDim.x * blockDim.x) {
    const unsigned node = list[j];
    const int x = (int)(node % (unsigned)b.nx);
    const int y = (int)((node / (unsigned)b.nx) % (unsigned)b.ny);
    const int z = (int)(node / ((unsigned)b.nx * (unsigned)b.ny));
    for (int i = 0; i < Q39; ++i) {
      const double v = (double)f[(size_t)i * n + node] + rf.f[i];
      double r = __builtin_nan("");
      if (v < 0.0) {
        const int cx = cx39<kShifted>(i), cy = CY39[i], cz = CZ39[i];
        // a neighbour across a non-periodic face counts as not positive: the
        // one-sided form of damp_it.f95:109-123
        const long long kp = damp_nb(x + cx, y + cy, z + cz, b, ghost);
        const long long km = damp_nb(x - cx, y - cy, z - cz, b, ghost);
        // with a link table (damp_it.f95:125-177): the + neighbour counts when
        // the link code is >= -500 (the rest direction: the node's code), the
        // - neighbour when its node code is >= -500; with neither, no repair
        // a plain-lattice table (damp_it.f95:64-123): the + neighbour counts
        // when the node's link code of i is >= 0, the - neighbour when that
        // of opp(i) is (the rest direction: the node's code); with neither, |f|
        bool pok = kp >= 0, mok = km >= 0;
        if (state && kShifted) {
          pok = pok && state[(size_t)(i == 0 ? 0 : i) * n + node] >= -500;
          mok = mok && state[(size_t)km] >= -500;
        } else if (state) {
          pok = pok && state[(size_t)i * n + node] >= 0;
          mok = mok && state[(size_t)OPP39[i] * n + node] >= 0;
        }
        const double fp = pok ? (double)f[(size_t)i * n + (size_t)kp] + rf.f[i] : -1.0;
        const double fm = mok ? (double)f[(size_t)i * n + (size_t)km] + rf.f[i] : -1.0;
        if (pok && mok) {
          if (fp > 0.0 && fm > 0.0) r = (fp + fm) / 2.0;
          else if (fp > 0.0 && fm < 0.0) r = (fp + fabs(v)) / 2;
          else if (fp < 0.0 && fm > 0.0) r = (fm + fabs(v)) / 2;
          else r = fabs(v);
        } else if (pok) {
          r = fp > 0.0 ? (fp + fabs(v)) / 2 : fabs(v);
        } else if (mok) {
          r = fm > 0.0 ? (fm + fabs(v)) / 2 : fabs(v);
        } else if (!(state && kShifted)) {
          r = fabs(v);
        }
        if (r == r) ++nrep;
      }
      rep[(size_t)j * Q39 + i] = r;
    }
  }
  if (nrep) atomicAdd(stats + (size_t)(blockIdx.x & (kStatSlots - 1)) * kStatFields + ST_REPAIR, nrep);
}

template <bool kFp32>
__global__ __launch_bounds__(256) void k39_damp_apply(void* __restrict__ f_, Box b, Ref39 rf,
                                                      unsigned* __restrict__ count,
                                                      const unsigned* __restrict__ list,
                                                      const double* __restrict__ rep) {
  using S = typename std::conditional<kFp32, float, double>::type;
  S* __restrict__ f = static_cast<S*>(f_);
  const unsigned cnt = *count;
  const unsigned n = (unsigned)b.n;
  for (unsigned j = blockIdx.x * blockDim.x + threadIdx.x; j < cnt; j += gridDim.x * blockDim.x) {
    const unsigned node = list[j];
    for (int i = 0; i < Q39; ++i) {
      const double r = rep[(size_t)j * Q39 + i];
      if (r == r) f[(size_t)i * n + node] = (S)(r - rf.f[i]);
    }
  }
}

__global__ void k39_clear(unsigned* c) { *c = 0u; }

// Equilibrium state of given rho, u, T per node: Newton from the Maxwellian's
// multipliers, f = f_eq, g = (2 cv - 3) T f_eq (init_fi.f95:41 sets the
// populations to the equilibrium of the initial state the same way).
template <bool kFp32, bool kShifted>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(1, 2))) void k39_init(
    const double* __restrict__ fields, void* __restrict__ f_, void* __restrict__ g_,
    double* __restrict__ lmv, Box b, Gas39 gs, Ref39 rf, unsigned long long* __restrict__ bad) {
  using S = typename std::conditional<kFp32, float, double>::type;
  S* __restrict__ fo = static_cast<S*>(f_);
  S* __restrict__ go = static_cast<S*>(g_);
  const unsigned n = (unsigned)b.n;
  const unsigned node = blockIdx.x * 64u + threadIdx.x;
  if (node >= n) return;
  const double rho = fields[node], u0 = fields[(size_t)n + node], u1 = fields[2 * (size_t)n + node],
               u2 = fields[3 * (size_t)n + node], T = fields[4 * (size_t)n + node];
  const double vm = u0 * u0 + u1 * u1 + u2 * u2, qf = rho * vm + 5 * rho * T;
  const double tg[kM39] = {rho, rho * u0, rho * u1, rho * u2, rho * (T + u0 * u0),
                           rho * (T + u1 * u1), rho * (T + u2 * u2), rho * u0 * u1, rho * u0 * u2,
                           rho * u1 * u2, qf * u0, qf * u1, qf * u2};
  double lm[kM39] = {-1.5 * log(2 * 3.141592653589793 * T) - vm / (2 * T), u0 / T, u1 / T, u2 / T,
                     -1 / (2 * T), -1 / (2 * T), -1 / (2 * T), 0, 0, 0, 0, 0, 0};
  double lmp[kM39];
  const int it = lm_newton39<kShifted>(rho, tg, lm, lmp, 200);
  if (it < 0) atomicAdd(bad, 1ull);
#pragma unroll
  for (int k = 0; k < kM39; ++k) lmv[(size_t)k * n + node] = lm[k];
#pragma unroll
  for (int i = 0; i < Q39; ++i) {
    const double e = expo39<kShifted>(lm, rho, i);
    fo[(size_t)i * n + node] = (S)(e - rf.f[i]);
    go[(size_t)i * n + node] = (S)(gs.lead_g * T * e - rf.g[i]);
  }
}

// Ghost nodes of a refined D3Q39 level from the coarser level
// (interlevel_interpolation.f95:114-493 through amrex_fillcoarsepatch with
// amrex_interp_node_bilinear): per axis the coincident coarse node or the mean
// of the two around (ratio 2, nodal), in time (1 - theta) * previous + theta *
// current, for f and g in their storage form (the reference state is the same
// on every level) and for rho and T of the coarse nodes.  A population with
// f or g negative or NaN becomes the equilibrium of the interpolated coarse
// state, f = rho exp(lambda . m_i), g = (2 cv - 3) T f (:417-431); the coarse
// level keeps no previous multipliers, so lambda is interpolated in space from
// its current ones.  The multipliers are stored too (warm start of a node that
// later joins the valid region).
template <bool kFp32, bool kShifted>
__global__ __launch_bounds__(64) void k39_coarse_to_fine(
    const void* __restrict__ cpf_, const void* __restrict__ cnf_, const void* __restrict__ cpg_,
    const void* __restrict__ cng_, const double* __restrict__ clm, LevelMap cm,
    void* __restrict__ ff_, void* __restrict__ fg_, double* __restrict__ flm, LevelMap fm,
    const unsigned* __restrict__ ghosts, unsigned nghost, double theta, Ref39 rf, Gas39 gs,
    unsigned long long* __restrict__ stats) {
  using S = typename std::conditional<kFp32, float, double>::type;
  const S* __restrict__ cpf = static_cast<const S*>(cpf_);
  const S* __restrict__ cnf = static_cast<const S*>(cnf_);
  const S* __restrict__ cpg = static_cast<const S*>(cpg_);
  const S* __restrict__ cng = static_cast<const S*>(cng_);
  S* __restrict__ ff = static_cast<S*>(ff_);
  S* __restrict__ fg = static_cast<S*>(fg_);
  const unsigned j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nghost) return;
  const unsigned node = ghosts ? ghosts[j] : j;
  const unsigned fnx = (unsigned)fm.box.nx, fny = (unsigned)fm.box.ny;
  const int loc[3] = {(int)(node % fnx), (int)((node / fnx) % fny), (int)(node / (fnx * fny))};
  const int cn[3] = {cm.box.nx, cm.box.ny, cm.box.nz};
  int ci[3][2];
  double wt[3][2];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const int G = fm.origin[d] + loc[d];
    const int odd = G & 1;
    int c0 = (G >> 1) - cm.origin[d];
    int c1 = c0 + odd;
    if (cm.per[d]) {
      c0 = (c0 % cn[d] + cn[d]) % cn[d];
      c1 = (c1 % cn[d] + cn[d]) % cn[d];
    }
    ci[d][0] = c0;
    ci[d][1] = c1;
    wt[d][0] = odd ? 0.5 : 1.0;
    wt[d][1] = odd ? 0.5 : 0.0;
  }
  const unsigned cnn = (unsigned)cm.box.n, fnn = (unsigned)fm.box.n;
  double f[Q39], g[Q39], lm[kM39];
#pragma unroll
  for (int i = 0; i < Q39; ++i) f[i] = g[i] = 0.0;
#pragma unroll
  for (int k = 0; k < kM39; ++k) lm[k] = 0.0;
  double rho_i = 0.0, T_i = 0.0;
  for (int cz = 0; cz < 2; ++cz)
    for (int cy = 0; cy < 2; ++cy)
      for (int cx = 0; cx < 2; ++cx) {
        const double w = wt[0][cx] * wt[1][cy] * wt[2][cz];
        if (w == 0.0) continue;
        const unsigned k = (unsigned)ci[0][cx] +
                           (unsigned)cm.box.nx * ((unsigned)ci[1][cy] + (unsigned)cm.box.ny * (unsigned)ci[2][cz]);
        // rho and T of the coarse node at both times (calculate_flow_variables.f95:201-228)
        double r[2] = {0.0, 0.0}, mx[2] = {0.0, 0.0}, my[2] = {0.0, 0.0}, mz[2] = {0.0, 0.0},
               en[2] = {0.0, 0.0};
#pragma unroll
        for (int i = 0; i < Q39; ++i) {
          const double ap = (double)cpf[(size_t)i * cnn + k] + rf.f[i];
          const double an = (double)cnf[(size_t)i * cnn + k] + rf.f[i];
          const double bp = (double)cpg[(size_t)i * cnn + k] + rf.g[i];
          const double bn = (double)cng[(size_t)i * cnn + k] + rf.g[i];
          f[i] += w * ((1.0 - theta) * (double)cpf[(size_t)i * cnn + k] + theta * (double)cnf[(size_t)i * cnn + k]);
          g[i] += w * ((1.0 - theta) * (double)cpg[(size_t)i * cnn + k] + theta * (double)cng[(size_t)i * cnn + k]);
          const int ccx = cx39<kShifted>(i), ccy = CY39[i], ccz = CZ39[i];
          const double cm2 = (double)(ccx * ccx + ccy * ccy + ccz * ccz);
          r[0] += ap; r[1] += an;
          mx[0] += ccx * ap; mx[1] += ccx * an;
          my[0] += ccy * ap; my[1] += ccy * an;
          mz[0] += ccz * ap; mz[1] += ccz * an;
          en[0] += cm2 * ap + bp; en[1] += cm2 * an + bn;
        }
        double Tn[2];
#pragma unroll
        for (int t = 0; t < 2; ++t) {
          const double ux = mx[t] / r[t], uy = my[t] / r[t], uz = mz[t] / r[t];
          Tn[t] = (en[t] - r[t] * (ux * ux + uy * uy + uz * uz)) / (2 * gs.cv * r[t]);
        }
        rho_i += w * ((1.0 - theta) * r[0] + theta * r[1]);
        T_i += w * ((1.0 - theta) * Tn[0] + theta * Tn[1]);
#pragma unroll
        for (int q = 0; q < kM39; ++q) lm[q] += w * clm[(size_t)q * cnn + k];
      }
  unsigned long long nrep = 0;
#pragma unroll
  for (int i = 0; i < Q39; ++i) {
    const double fv = f[i] + rf.f[i], gv = g[i] + rf.g[i];
    if (!(fv >= 0.0) || !(gv >= 0.0)) {
      const double e = expo39<kShifted>(lm, rho_i, i);
      f[i] = e - rf.f[i];
      g[i] = gs.lead_g * T_i * e - rf.g[i];
      ++nrep;
    }
    ff[(size_t)i * fnn + node] = (S)f[i];
    fg[(size_t)i * fnn + node] = (S)g[i];
  }
#pragma unroll
  for (int q = 0; q < kM39; ++q) flm[(size_t)q * fnn + node] = lm[q];
  if (nrep) atomicAdd(stats + (size_t)(blockIdx.x & (kStatSlots - 1)) * kStatFields + ST_REPAIR, nrep);
}

// fine_to_coarse.f95:25-65 (amrex_average_down_nodes at ratio 2: the
// coincident fine node) for f, g and the multipliers of the covered coarse
// nodes; both levels hold the same storage form
template <typename S>
__global__ __launch_bounds__(256) void k39_fine_to_coarse(
    const S* __restrict__ ff, const S* __restrict__ fg, const double* __restrict__ flm, LevelMap fm,
    S* __restrict__ cf, S* __restrict__ cg, double* __restrict__ clm, LevelMap cm, Range3 cover) {
  const int ex = cover.hi[0] - cover.lo[0] + 1, ey = cover.hi[1] - cover.lo[1] + 1;
  const int ez = cover.hi[2] - cover.lo[2] + 1;
  const unsigned j = blockIdx.x * blockDim.x + threadIdx.x;
  if (ex <= 0 || ey <= 0 || ez <= 0 || j >= (unsigned)(ex * ey * ez)) return;
  const int C[3] = {cover.lo[0] + (int)(j % (unsigned)ex), cover.lo[1] + (int)((j / (unsigned)ex) % (unsigned)ey),
                    cover.lo[2] + (int)(j / ((unsigned)ex * (unsigned)ey))};
  int cl[3], fl[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    cl[d] = C[d] - cm.origin[d];
    fl[d] = 2 * C[d] - fm.origin[d];
  }
  const unsigned ck = (unsigned)cl[0] + (unsigned)cm.box.nx * ((unsigned)cl[1] + (unsigned)cm.box.ny * (unsigned)cl[2]);
  const unsigned fk = (unsigned)fl[0] + (unsigned)fm.box.nx * ((unsigned)fl[1] + (unsigned)fm.box.ny * (unsigned)fl[2]);
  const size_t cnn = (size_t)cm.box.n, fnn = (size_t)fm.box.n;
  for (int i = 0; i < Q39; ++i) {
    cf[i * cnn + ck] = ff[i * fnn + fk];
    cg[i * cnn + ck] = fg[i * fnn + fk];
  }
  for (int q = 0; q < kM39; ++q) clm[q * cnn + ck] = flm[q * fnn + fk];
}

// host layout (FP64 full values) <-> device storage
template <bool kFp32>
__global__ void k39_from_fp64(const double* __restrict__ src, void* __restrict__ dst_,
                              const double* __restrict__ ref, size_t n) {
  using S = typename std::conditional<kFp32, float, double>::type;
  S* dst = static_cast<S*>(dst_);
  for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < n * Q39;
       k += (size_t)gridDim.x * blockDim.x)
    dst[k] = (S)(src[k] - ref[k / n]);
}

// flow variables of calculate_flow_variables.f95:201-221 at stored node k of a
// level of n nodes: rho, u, v, w, T and the reference's press = rho (T + |u|^2 / 2)
template <bool kFp32, bool kShifted>
__device__ __forceinline__ void flow39(const void* __restrict__ f_, const void* __restrict__ g_,
                                       size_t n, size_t k, const Ref39& rf, double cv,
                                       double o[6]) {
  using S = typename std::conditional<kFp32, float, double>::type;
  const S* __restrict__ fA = static_cast<const S*>(f_);
  const S* __restrict__ gA = static_cast<const S*>(g_);
  double r = 0.0, mx = 0.0, my = 0.0, mz = 0.0, en = 0.0;
#pragma unroll
  for (int i = 0; i < Q39; ++i) {
    const int cx = cx39<kShifted>(i), cy = CY39[i], cz = CZ39[i];
    const double f = (double)fA[(size_t)i * n + k] + rf.f[i],
                 g = (double)gA[(size_t)i * n + k] + rf.g[i];
    r += f;
    mx += (double)cx * f;
    my += (double)cy * f;
    mz += (double)cz * f;
    en += (double)(cx * cx + cy * cy + cz * cz) * f + g;
  }
  const double u = mx / r, v = my / r, w = mz / r, vm = u * u + v * v + w * w;
  const double T = (en - r * vm) / (2 * cv * r);
  o[0] = r;
  o[1] = u;
  o[2] = v;
  o[3] = w;
  o[4] = T;
  o[5] = r * (T + vm / 2.0);
}

// every node: out is [6][n]
template <bool kFp32, bool kShifted>
__global__ __launch_bounds__(256) void k39_macro(const void* __restrict__ f_,
                                                 const void* __restrict__ g_, Box b, Ref39 rf,
                                                 double cv, double* __restrict__ out) {
  const size_t n = (size_t)b.n;
  for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < n;
       k += (size_t)gridDim.x * blockDim.x) {
    double o[6];
    flow39<kFp32, kShifted>(f_, g_, n, k, rf, cv, o);
#pragma unroll
    for (int j = 0; j < 6; ++j) out[j * n + k] = o[j];
  }
}

// listed stored nodes: out is [6][m]
template <bool kFp32, bool kShifted>
__global__ __launch_bounds__(64) void k39_probe(const void* __restrict__ f_,
                                                const void* __restrict__ g_, size_t n,
                                                const uint32_t* __restrict__ nodes, unsigned m,
                                                Ref39 rf, double cv, double* __restrict__ out) {
  const unsigned j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= m) return;
  double o[6];
  flow39<kFp32, kShifted>(f_, g_, n, nodes[j], rf, cv, o);
#pragma unroll
  for (int c = 0; c < 6; ++c) out[(size_t)c * m + j] = o[c];
}

template <bool kFp32>
__global__ void k39_to_fp64(const void* __restrict__ src_, double* __restrict__ dst,
                            const double* __restrict__ ref, size_t n) {
  using S = typename std::conditional<kFp32, float, double>::type;
  const S* src = static_cast<const S*>(src_);
  for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < n * Q39;
       k += (size_t)gridDim.x * blockDim.x)
    dst[k] = (double)src[k] + ref[k / n];
}

}  // namespace

hipError_t launch_d3q39_step(const Box& b, const Step39& a, hipStream_t s) {
  const unsigned blocks = (unsigned)((b.n + 63) / 64);
  const unsigned near_blocks = (a.n_near + 63u) / 64u;
  if (a.has_faces && a.n_near && !a.near_list) return hipErrorInvalidValue;
  const unsigned wall_blocks = (a.walls.n + 63u) / 64u;
  hipLaunchKernelGGL(k39_clear, dim3(1), dim3(1), 0, s, a.neg_count);
  // the wall nodes' rebuilt populations, from the input buffers
  if (a.walls.cls && a.walls.n) {
#define UCLBS_WALLS39(K)                                                                         \
  hipLaunchKernelGGL(K, dim3(wall_blocks), dim3(64), 0, s, a.fin, a.gin, a.lm, b, a.gas, a.ref, \
                     a.faces, (long long)a.step, a.walls, a.stats)
    if (a.shifted) {
      if (a.fp32) UCLBS_WALLS39(k39_wall_rebuild<true>); else UCLBS_WALLS39(k39_wall_rebuild<false>);
    } else {
      if (a.fp32) UCLBS_WALLS39(k39_wall_rebuild_plain<true>);
      else UCLBS_WALLS39(k39_wall_rebuild_plain<false>);
    }
#undef UCLBS_WALLS39
  }
#define UCLBS_K39(F, SH, TU, KA, NR, NB)                                                         \
  hipLaunchKernelGGL((k39_step<F, SH, TU, KA, NR>), dim3(NB), dim3(64), 0, s, a.fin, a.gin,     \
                     a.fout, a.gout, a.lm, a.alpha, b, a.gas, a.ref, (long long)a.step, a.stats, \
                     a.neg_count, a.neg_list, a.faces, a.has_faces, a.near_list, a.n_near,   \
                     a.ghost, a.walls)
#define UCLBS_K39_FC(F, SH, TU, KA)                                              \
  do {                                                                           \
    UCLBS_K39(F, SH, TU, KA, 0, blocks);                                         \
    if (a.has_faces && a.n_near) UCLBS_K39(F, SH, TU, KA, 1, near_blocks);       \
    if (a.walls.cls && a.walls.n) UCLBS_K39(F, SH, TU, KA, 2, wall_blocks);      \
  } while (0)
#define UCLBS_K39_KA(F, SH, TU) \
  do {                          \
    if (a.alpha) UCLBS_K39_FC(F, SH, TU, true); else UCLBS_K39_FC(F, SH, TU, false); \
  } while (0)
#define UCLBS_K39_TU(F, SH) \
  do {                      \
    if (a.gas.turbulent) UCLBS_K39_KA(F, SH, true); else UCLBS_K39_KA(F, SH, false); \
  } while (0)
  if (a.fp32) {
    if (a.shifted) UCLBS_K39_TU(true, true); else UCLBS_K39_TU(true, false);
  } else {
    if (a.shifted) UCLBS_K39_TU(false, true); else UCLBS_K39_TU(false, false);
  }
#undef UCLBS_K39_TU
#undef UCLBS_K39_KA
#undef UCLBS_K39_FC
#undef UCLBS_K39
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  // repair pass over the nodes that produced a negative population; the
  // grids read the count on the device, so no host synchronisation is needed
  const dim3 rg(256), rb(256);
  if (a.fp32) {
    if (a.shifted)
      hipLaunchKernelGGL((k39_damp_compute<true, true>), rg, rb, 0, s, a.fout, b, a.ref,
                         a.neg_count, a.neg_list, a.rep, a.stats, a.ghost, a.walls.state);
    else
      hipLaunchKernelGGL((k39_damp_compute<true, false>), rg, rb, 0, s, a.fout, b, a.ref,
                         a.neg_count, a.neg_list, a.rep, a.stats, a.ghost, a.walls.state);
    hipLaunchKernelGGL((k39_damp_apply<true>), rg, rb, 0, s, a.fout, b, a.ref, a.neg_count,
                       a.neg_list, a.rep);
  } else {
    if (a.shifted)
      hipLaunchKernelGGL((k39_damp_compute<false, true>), rg, rb, 0, s, a.fout, b, a.ref,
                         a.neg_count, a.neg_list, a.rep, a.stats, a.ghost, a.walls.state);
    else
      hipLaunchKernelGGL((k39_damp_compute<false, false>), rg, rb, 0, s, a.fout, b, a.ref,
                         a.neg_count, a.neg_list, a.rep, a.stats, a.ghost, a.walls.state);
    hipLaunchKernelGGL((k39_damp_apply<false>), rg, rb, 0, s, a.fout, b, a.ref, a.neg_count,
                       a.neg_list, a.rep);
  }
  return hipGetLastError();
}

hipError_t launch_d3q39_init(const Box& b, const double* fields, void* f, void* g, double* lm,
                             int fp32, int shifted, const Gas39& gs, const Ref39& rf,
                             unsigned long long* bad, hipStream_t s) {
  const unsigned blocks = (unsigned)((b.n + 63) / 64);
  if (fp32) {
    if (shifted)
      hipLaunchKernelGGL((k39_init<true, true>), dim3(blocks), dim3(64), 0, s, fields, f, g, lm, b,
                         gs, rf, bad);
    else
      hipLaunchKernelGGL((k39_init<true, false>), dim3(blocks), dim3(64), 0, s, fields, f, g, lm,
                         b, gs, rf, bad);
  } else {
    if (shifted)
      hipLaunchKernelGGL((k39_init<false, true>), dim3(blocks), dim3(64), 0, s, fields, f, g, lm,
                         b, gs, rf, bad);
    else
      hipLaunchKernelGGL((k39_init<false, false>), dim3(blocks), dim3(64), 0, s, fields, f, g, lm,
                         b, gs, rf, bad);
  }
  return hipGetLastError();
}

hipError_t launch_d3q39_from_fp64(const Box& b, const double* src, void* dst, int fp32,
                                  const double* ref, hipStream_t s) {
  const size_t tot = (size_t)b.n * Q39;
  const unsigned blocks = (unsigned)std::min<size_t>((tot + 255) / 256, 65536);
  if (fp32) hipLaunchKernelGGL((k39_from_fp64<true>), dim3(blocks), dim3(256), 0, s, src, dst, ref, (size_t)b.n);
  else hipLaunchKernelGGL((k39_from_fp64<false>), dim3(blocks), dim3(256), 0, s, src, dst, ref, (size_t)b.n);
  return hipGetLastError();
}

hipError_t launch_d3q39_to_fp64(const Box& b, const void* src, double* dst, int fp32,
                                const double* ref, hipStream_t s) {
  const size_t tot = (size_t)b.n * Q39;
  const unsigned blocks = (unsigned)std::min<size_t>((tot + 255) / 256, 65536);
  if (fp32) hipLaunchKernelGGL((k39_to_fp64<true>), dim3(blocks), dim3(256), 0, s, src, dst, ref, (size_t)b.n);
  else hipLaunchKernelGGL((k39_to_fp64<false>), dim3(blocks), dim3(256), 0, s, src, dst, ref, (size_t)b.n);
  return hipGetLastError();
}

hipError_t launch_d3q39_macro(const Box& b, const void* f, const void* g, int fp32, int shifted,
                              const Ref39& rf, double cv, double* out, hipStream_t s) {
  const unsigned blocks = (unsigned)std::min<size_t>(((size_t)b.n + 255) / 256, 65536);
  if (fp32) {
    if (shifted) hipLaunchKernelGGL((k39_macro<true, true>), dim3(blocks), dim3(256), 0, s, f, g, b, rf, cv, out);
    else hipLaunchKernelGGL((k39_macro<true, false>), dim3(blocks), dim3(256), 0, s, f, g, b, rf, cv, out);
  } else {
    if (shifted) hipLaunchKernelGGL((k39_macro<false, true>), dim3(blocks), dim3(256), 0, s, f, g, b, rf, cv, out);
    else hipLaunchKernelGGL((k39_macro<false, false>), dim3(blocks), dim3(256), 0, s, f, g, b, rf, cv, out);
  }
  return hipGetLastError();
}

hipError_t launch_d3q39_probe(const void* f, const void* g, int64_t n, int fp32, int shifted,
                              const Ref39& rf, double cv, const uint32_t* nodes, unsigned m,
                              double* out, hipStream_t s) {
  if (m == 0) return hipSuccess;
  const dim3 grid((m + 63) / 64), blk(64);
  if (fp32) {
    if (shifted) hipLaunchKernelGGL((k39_probe<true, true>), grid, blk, 0, s, f, g, (size_t)n, nodes, m, rf, cv, out);
    else hipLaunchKernelGGL((k39_probe<true, false>), grid, blk, 0, s, f, g, (size_t)n, nodes, m, rf, cv, out);
  } else {
    if (shifted) hipLaunchKernelGGL((k39_probe<false, true>), grid, blk, 0, s, f, g, (size_t)n, nodes, m, rf, cv, out);
    else hipLaunchKernelGGL((k39_probe<false, false>), grid, blk, 0, s, f, g, (size_t)n, nodes, m, rf, cv, out);
  }
  return hipGetLastError();
}

hipError_t launch_coarse_to_fine39(const void* cpf, const void* cnf, const void* cpg,
                                   const void* cng, const double* clm, const LevelMap& cm, void* ff,
                                   void* fg, double* flm, const LevelMap& fm,
                                   const unsigned* ghosts, unsigned nghost, double theta, int fp32,
                                   int shifted, const Ref39& rf, const Gas39& gs,
                                   unsigned long long* stats, hipStream_t s) {
  if (nghost == 0) return hipSuccess;
  const dim3 g((nghost + 63) / 64), b(64);
#define UCLBS_C2F39(F, SH)                                                                      \
  hipLaunchKernelGGL((k39_coarse_to_fine<F, SH>), g, b, 0, s, cpf, cnf, cpg, cng, clm, cm, ff, fg, \
                     flm, fm, ghosts, nghost, theta, rf, gs, stats)
  if (fp32) {
    if (shifted) UCLBS_C2F39(true, true); else UCLBS_C2F39(true, false);
  } else {
    if (shifted) UCLBS_C2F39(false, true); else UCLBS_C2F39(false, false);
  }
#undef UCLBS_C2F39
  return hipGetLastError();
}

hipError_t launch_fine_to_coarse39(const void* ff, const void* fg, const double* flm,
                                   const LevelMap& fm, void* cf, void* cg, double* clm,
                                   const LevelMap& cm, const Range3& cover, int fp32, hipStream_t s) {
  const long long cnt = (long long)(cover.hi[0] - cover.lo[0] + 1) * (cover.hi[1] - cover.lo[1] + 1) *
                        (cover.hi[2] - cover.lo[2] + 1);
  if (cnt <= 0) return hipSuccess;
  const dim3 g((unsigned)((cnt + 255) / 256)), b(256);
  if (fp32)
    hipLaunchKernelGGL((k39_fine_to_coarse<float>), g, b, 0, s, static_cast<const float*>(ff),
                       static_cast<const float*>(fg), flm, fm, static_cast<float*>(cf),
                       static_cast<float*>(cg), clm, cm, cover);
  else
    hipLaunchKernelGGL((k39_fine_to_coarse<double>), g, b, 0, s, static_cast<const double*>(ff),
                       static_cast<const double*>(fg), flm, fm, static_cast<double*>(cf),
                       static_cast<double*>(cg), clm, cm, cover);
  return hipGetLastError();
}

}  // namespace uclbs
