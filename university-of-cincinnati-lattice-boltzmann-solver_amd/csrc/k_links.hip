// Device-side link / neighbour table builder (compiled with -ffp-contract=off).
//
// Builds the reference's `state` table, state[q*n + node] in int32, for one
// nodal box directly in HBM:
//   component 0, node code: lbgrid/fluid_or_solid.f95:86-442.  A ray from the
//     node away from the grid centre (utilities/ray_casting_away.f95), length
//     100 x the largest extent, box-tested (utilities/box_test.f95), cut with
//     each triangle plane (fluid_or_solid.f95:176-196) and sign-tested
//     (utilities/tri_volume.f95); odd parity = solid (-1001), even = fluid (0);
//     one odd edge hit with even crossings counts as a crossing (:393-399).
//   components 1..26, link codes: lbgrid/state_directions.f95 -- a link to a
//     solid node gets the inter_bb_setup.f95:44-246 code
//     -1001 - floor(|crossing - node| / (|c| dx) * 1e9) of the first triangle
//     (index order) the segment crosses, or 0 when none is crossed; a link
//     leaving the box takes the face code (faces in the order x-, y-, z-, x+,
//     y+, z+, :194-684) unless the face is periodic; any other link takes the
//     neighbour's node code (:687-716); a solid node gets -666 everywhere (:87).
// Ties that make the reference re-cast with a time-seeded random ray
// (utilities/ray_casting.f95) are re-cast here along three fixed directions
// and counted.  One thread per node; the triangle soup and its bounding boxes
// are read in index order by every thread (the same addresses across a
// wavefront, so they stream through the cache).  The arithmetic is written in
// the reference's expression order with contraction off, so the table is
// bit-identical to the CPU restatement.
#pragma clang fp contract(off)
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <vector>

#include "kernels.h"
#include "lattice.h"

namespace uclbs {
namespace {

struct FaceCodes {
  int32_t code[6];
};

constexpr double kPlusEps = 1.24e-16;  // general/modules/precise.f95:13
constexpr double kMinusEps = -1.24e-16;
constexpr double kVSmall = 1.0e-14;    // precise.f95:16

__device__ __forceinline__ bool box_out(const double* p1, const double* p2, const double* bb) {
  for (int d = 0; d < 3; ++d) {
    const double mx = bb[2 * d], mn = bb[2 * d + 1];
    if (p1[d] >= mx + kPlusEps && p2[d] >= mx + kPlusEps) return true;
    if (p1[d] <= mn + kMinusEps && p2[d] <= mn + kMinusEps) return true;
  }
  return false;
}

__device__ __forceinline__ int tri_volume(const double* xc, const double* e, const double* t) {
  const double x1 = t[0] - e[0], x2 = t[3] - e[0], x3 = t[6] - e[0];
  const double y1 = t[1] - e[1], y2 = t[4] - e[1], y3 = t[7] - e[1];
  const double z1 = t[2] - e[2], z2 = t[5] - e[2], z3 = t[8] - e[2];
  const double px = xc[0] - e[0], py = xc[1] - e[1], pz = xc[2] - e[2];
  double v[3];
  v[0] = px * (y1 * z2 - z1 * y2) + py * (z1 * x2 - x1 * z2) + pz * (x1 * y2 - y1 * x2);
  v[1] = -px * (y1 * z3 - z1 * y3) - py * (z1 * x3 - x1 * z3) - pz * (x1 * y3 - y1 * x3);
  v[2] = px * (y2 * z3 - z2 * y3) + py * (z2 * x3 - x2 * z3) + pz * (x2 * y3 - y2 * x3);
  int zc = 0;
  for (int i = 0; i < 3; ++i)
    if (v[i] >= kMinusEps && v[i] <= kPlusEps) zc++;
  if (v[0] > 0.0 && v[1] > 0.0 && v[2] > 0.0) return 1;
  if (v[0] < 0.0 && v[1] < 0.0 && v[2] < 0.0) return 1;
  if (zc == 1) return -1;
  if (zc == 2) return -2;
  if (zc == 3) return -3;
  return 0;
}

// Candidate triangles of a segment: the triangles whose box box_out does not
// cull, found through the hierarchy.  A node whose box (the exact union of its
// triangles' boxes) box_out culls holds only triangles it culls one by one
// (the comparisons are monotone in the box bounds), so the candidates are the
// same set the reference's scan over every triangle keeps; visit(i) is called
// for each, in no particular order, and stops the walk by returning false.
template <class Visit>
__device__ __forceinline__ void for_candidates(const LinkGeom& g, const double* e1,
                                               const double* e2, Visit visit) {
  if (g.ntri <= 0) return;
  int stack[40];
  int sp = 0;
  stack[sp++] = 0;
  while (sp > 0) {
    const BvhNode& nd = g.bvh[stack[--sp]];
    if (box_out(e1, e2, nd.bb)) continue;
    if (nd.b < 0) {
      for (int j = 0; j < -nd.b; ++j) {
        const int i = g.bvh_idx[nd.a + j];
        if (box_out(e1, e2, g.tbb + 6 * (size_t)i)) continue;
        if (!visit(i)) return;
      }
    } else {
      stack[sp++] = nd.b;
      stack[sp++] = nd.a;
    }
  }
}

// parity of one cast: 0 even, 1 odd, -1 tie.  The parity counts do not depend
// on the order of the triangles and any tie gives -1, so the walk order of the
// hierarchy gives the reference's result.
__device__ int cast_parity(const LinkGeom& g, const double* e1, const double* e2) {
  int cross = 0, edge = 0;
  bool tie = false;
  for_candidates(g, e1, e2, [&](int i) {
    const double* t = g.tri + 9 * (size_t)i;
    const double* n = g.nrm + 3 * (size_t)i;
    const double dot = n[0] * t[0] + n[1] * t[1] + n[2] * t[2];
    const double upper = dot - (e1[0] * n[0] + e1[1] * n[1] + e1[2] * n[2]);
    const double lower =
        (e2[0] - e1[0]) * n[0] + (e2[1] - e1[1]) * n[1] + (e2[2] - e1[2]) * n[2];
    if (lower >= kMinusEps && lower <= kPlusEps) { tie = true; return false; }
    const double pc = upper / lower;
    if (pc > 0.0 && pc < 1.0) {
      const double xc[3] = {e1[0] + pc * (e2[0] - e1[0]), e1[1] + pc * (e2[1] - e1[1]),
                            e1[2] + pc * (e2[2] - e1[2])};
      if (upper >= lower + kMinusEps && upper <= lower + kPlusEps) { tie = true; return false; }
      const int v = tri_volume(xc, e1, t);
      if (v == 1) cross++;
      else if (v == -1) edge++;
      else if (v <= -2) { tie = true; return false; }
    } else if (upper >= -kVSmall && upper <= kVSmall) {
      tie = true;
      return false;
    }
    return true;
  });
  if (tie) return -1;
  if (edge >= 2) return -1;
  if ((edge % 2) == 1 && (cross % 2) == 0) { edge = 0; cross++; }
  edge /= 2;
  return (cross + edge) % 2;
}

__constant__ double kRecast[3][3] = {{0.5773502691896258, 0.6311775151936213, 0.5179036655461733},
                                     {-0.7071067811865476, 0.3826834323650898, 0.5946035575013605},
                                     {0.2672612419124244, -0.5345224838248488, 0.8017837257372732}};

// fluid (0) or solid (1) of the node at index position (a, b, c) -- inside the
// box or not -- by the ray cast of fluid_or_solid.f95; *tie set on a re-cast
__device__ int classify_pos(const LinkGeom& g, int a, int b, int c, int* tie) {
  const double e1[3] = {g.origin[0] + (double)a * g.dx, g.origin[1] + (double)b * g.dx,
                        g.origin[2] + (double)c * g.dx};
  double cen[3], ext = 0.0;
  const int nd[3] = {g.box.nx, g.box.ny, g.box.nz};
  for (int d = 0; d < 3; ++d) {
    const double lo = g.origin[d], hi = g.origin[d] + (double)(nd[d] - 1) * g.dx;
    cen[d] = (lo + hi) / 2.0;
    if (hi - lo > ext) ext = hi - lo;
  }
  const double L = fabs(100.0 * ext);
  const double dl[3] = {e1[0] - cen[0], e1[1] - cen[1], e1[2] - cen[2]};
  const double dist = sqrt(dl[0] * dl[0] + dl[1] * dl[1] + dl[2] * dl[2]);
  int solid = 0;
  *tie = 0;
  if (g.ntri > 0) {
    for (int attempt = 0; attempt < 4; ++attempt) {
      double e2[3];
      if (attempt == 0 && g.ray_len > 0.0) {
        for (int d = 0; d < 3; ++d) e2[d] = e1[d] + g.ray_len * g.ray[d];
      } else if (attempt == 0 && dist > 0.0) {
        for (int d = 0; d < 3; ++d) e2[d] = e1[d] + L * dl[d] / dist;
      } else {
        const int r = attempt == 0 ? 0 : attempt - 1;
        const double Lr = g.ray_len > 0.0 ? g.ray_len : L;
        for (int d = 0; d < 3; ++d) e2[d] = e1[d] + Lr * kRecast[r][d];
        if (attempt == 0) *tie = 1;
      }
      if (box_out(e1, e2, g.gbb)) { solid = 0; break; }
      const int p = cast_parity(g, e1, e2);
      if (p >= 0) { solid = p; break; }
      *tie = 1;
    }
  }
  return solid;
}

__global__ __launch_bounds__(256) void k_classify(LinkGeom g, int32_t* __restrict__ state,
                                                  unsigned long long* __restrict__ counts) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= g.box.n) return;
  const int a = (int)(k % g.box.nx);
  const int64_t yz = k / g.box.nx;
  const int b = (int)(yz % g.box.ny), c = (int)(yz / g.box.ny);
  int tie = 0;
  const int solid = classify_pos(g, a, b, c, &tie);
  state[k] = solid ? -1001 : 0;
  if (solid) atomicAdd(counts + 0, 1ull);
  if (tie) atomicAdd(counts + 1, 1ull);
}

// Wall code of the link from node (a, b, c) along (cx, cy, cz): the reference
// takes the first triangle in index order that the segment crosses
// (inter_bb_setup.f95:44-246), flagging a tie on any earlier candidate whose
// plane holds the node.  Over the hierarchy's candidates: the crossing
// triangle of least index, and a tie when a tie candidate has a lower index
// or the crossing itself is one.
__device__ int32_t wall_code(const LinkGeom& g, int a, int b, int c, int cx, int cy, int cz,
                             int* tie) {
  const double e1[3] = {g.origin[0] + (double)a * g.dx, g.origin[1] + (double)b * g.dx,
                        g.origin[2] + (double)c * g.dx};
  const double e2[3] = {e1[0] + (double)cx * g.dx, e1[1] + (double)cy * g.dx,
                        e1[2] + (double)cz * g.dx};
  const double link = sqrt((double)(cx * cx + cy * cy + cz * cz)) * g.dx;
  int best = 0x7fffffff, tie_min = 0x7fffffff;
  bool best_tie = false;
  int32_t best_code = 0;
  for_candidates(g, e1, e2, [&](int i) {
    if (i > best) return true;  // a later triangle changes nothing
    const double* t = g.tri + 9 * (size_t)i;
    const double* n = g.nrm + 3 * (size_t)i;
    const double dot = n[0] * t[0] + n[1] * t[1] + n[2] * t[2];
    const double upper = dot - (e1[0] * n[0] + e1[1] * n[1] + e1[2] * n[2]);
    const double lower =
        (e2[0] - e1[0]) * n[0] + (e2[1] - e1[1]) * n[1] + (e2[2] - e1[2]) * n[2];
    if (lower >= kMinusEps && lower <= kPlusEps) return true;
    const double pc = upper / lower;
    if (pc > 0.0 && pc < 1.0 + kPlusEps) {
      const double xc[3] = {e1[0] + pc * (e2[0] - e1[0]), e1[1] + pc * (e2[1] - e1[1]),
                            e1[2] + pc * (e2[2] - e1[2])};
      bool hit, htie = false;
      if (upper >= lower - kVSmall && upper <= lower + kVSmall) { hit = true; htie = true; }
      else hit = tri_volume(xc, e1, t) != 0;
      if (hit) {
        const double d = sqrt((e1[0] - xc[0]) * (e1[0] - xc[0]) +
                              (e1[1] - xc[1]) * (e1[1] - xc[1]) +
                              (e1[2] - xc[2]) * (e1[2] - xc[2]));
        best = i;
        best_tie = htie;
        best_code = (int32_t)(-1001 - (int64_t)floor(d / link * 1e9));
      }
    } else if (upper >= -kVSmall && upper <= kVSmall) {
      if (i < tie_min) tie_min = i;
    }
    return true;
  });
  if (tie_min < best || (best != 0x7fffffff && best_tie)) *tie = 1;
  return best == 0x7fffffff ? 0 : best_code;
}

__global__ __launch_bounds__(256) void k_links(LinkGeom g, int32_t* __restrict__ state,
                                               unsigned long long* __restrict__ counts) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= g.box.n) return;
  const int a = (int)(k % g.box.nx);
  const int64_t yz = k / g.box.nx;
  const int b = (int)(yz % g.box.ny), c = (int)(yz / g.box.ny);
  const int64_t n = g.box.n;
  const int32_t own = state[k];
  int ties = 0;
  for (int i = 1; i < 27; ++i) {
    int32_t code;
    if (own == -1001 || own == -666) {
      code = -666;
    } else {
      const int cx = CX27[i], cy = CY27[i], cz = CZ27[i];
      int na = a + cx, nb = b + cy, nc = c + cz;
      const bool out[6] = {na < 0, nb < 0, nc < 0, na > g.box.nx - 1, nb > g.box.ny - 1,
                           nc > g.box.nz - 1};
      int face = -1;
      for (int f = 0; f < 6; ++f)
        if (out[f] && g.face_code[f] != kFacePeriodic) { face = f; break; }
      if (face >= 0) {
        code = g.face_code[face];
      } else {
        na = (na + g.box.nx) % g.box.nx;
        nb = (nb + g.box.ny) % g.box.ny;
        nc = (nc + g.box.nz) % g.box.nz;
        const int32_t nbr = state[na + (int64_t)g.box.nx * (nb + (int64_t)g.box.ny * nc)];
        if (nbr == -1001) {
          int tie = 0;
          code = wall_code(g, a, b, c, cx, cy, cz, &tie);
          ties += tie;
        } else {
          code = nbr;
        }
      }
    }
    state[(int64_t)i * n + k] = code;
  }
  if (ties) atomicAdd(counts + 1, (unsigned long long)ties);
}

// D3Q39 table on the plain lattice (lbgrid/state_directions.f95): one thread
// per node sets its links in the reference's direction order, so a long link
// reads the final codes of its shorter links (:95-185): (s + 1000) / 2 for the
// (+-2,0,0)-type from the speed-1 wall link, (s1 + 1000) / 3 or
// ((s2 + 1000) * 2) / 3 for the (+-3,0,0)-type -- the (0,3,0) link reading its
// own slot, as the reference reads state(...,35); the table is cleared first.
__global__ __launch_bounds__(256) void k_links39(LinkGeom g, int32_t* __restrict__ state,
                                                 unsigned long long* __restrict__ counts) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= g.box.n) return;
  const int a = (int)(k % g.box.nx);
  const int64_t yz = k / g.box.nx;
  const int b = (int)(yz % g.box.ny), c = (int)(yz / g.box.ny);
  const int64_t n = g.box.n;
  const int32_t own = state[k];
  int ties = 0;
  for (int i = 1; i < 39; ++i) {
    int32_t* slot = state + (int64_t)i * n + k;
    if (own == -1001 || own == -666) {
      *slot = -666;
      continue;
    }
    if (i >= 15 && i <= 20) {
      const int32_t s1 = state[(int64_t)(i - 14) * n + k];
      if (s1 < -1000) {
        *slot = (s1 + 1000) / 2;
        continue;
      }
    } else if (i >= 33) {
      const int32_t s1 = state[(int64_t)(i - 32) * n + k];
      const int32_t s2 = state[(int64_t)(i - 18) * n + k];
      if (s1 < -1000) {
        *slot = (s1 + 1000) / 3;
        continue;
      }
      if (s2 < -1000) {
        const int32_t v = (i == 34) ? *slot : s2;
        *slot = ((v + 1000) * 2) / 3;
        continue;
      }
    }
    const int cx = CX39[i], cy = CY39[i], cz = CZ39[i];
    int na = a + cx, nb = b + cy, nc = c + cz;
    const bool out[6] = {na < 0, nb < 0, nc < 0, na > g.box.nx - 1, nb > g.box.ny - 1,
                         nc > g.box.nz - 1};
    int face = -1;
    for (int f = 0; f < 6; ++f)
      if (out[f] && g.face_code[f] != kFacePeriodic) { face = f; break; }
    int32_t code;
    if (face >= 0) {
      code = g.face_code[face];
    } else {
      na = (na + g.box.nx) % g.box.nx;
      nb = (nb + g.box.ny) % g.box.ny;
      nc = (nc + g.box.nz) % g.box.nz;
      const int32_t nbr = state[na + (int64_t)g.box.nx * (nb + (int64_t)g.box.ny * nc)];
      if (nbr == -1001) {
        int tie = 0;
        code = wall_code(g, a, b, c, cx, cy, cz, &tie);
        ties += tie;
      } else {
        code = nbr;
      }
    }
    *slot = code;
  }
  if (ties) atomicAdd(counts + 1, (unsigned long long)ties);
}

// D3Q39 table on the shifted lattice (lbgrid/shifted_state_directions.f95):
// every link i of a fluid node x is keyed on its upstream node u = x - c_i
// (c_x shifted by +1): the wall code of the segment x -> u when u is solid
// (inside the box, or outside it by the same ray cast), else the code of the
// first non-periodic face u lies beyond (x-, y-, z-, x+, y+, z+; an outflow
// face, -6, gives the downstream node's code), else the downstream node
// x + c_i's code; -1001 becomes 0 (a solid downstream node, or a segment that
// crosses no triangle).  No link is derived from a shorter one.
__global__ __launch_bounds__(256) void k_links39s(LinkGeom g, int32_t* __restrict__ state,
                                                  unsigned long long* __restrict__ counts) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= g.box.n) return;
  const int nd[3] = {g.box.nx, g.box.ny, g.box.nz};
  const int x[3] = {(int)(k % nd[0]), (int)((k / nd[0]) % nd[1]), (int)(k / ((int64_t)nd[0] * nd[1]))};
  const int64_t n = g.box.n;
  const int32_t own = state[k];
  int ties = 0;
  for (int i = 1; i < 39; ++i) {
    int32_t* slot = state + (int64_t)i * n + k;
    if (own == -1001 || own == -666) {
      *slot = -666;
      continue;
    }
    const int cv[3] = {CX39[i] + 1, CY39[i], CZ39[i]};
    int up[3], dn[3];
    for (int d = 0; d < 3; ++d) {
      up[d] = x[d] - cv[d];
      dn[d] = x[d] + cv[d];
    }
    int face = -1;
    for (int f = 0; f < 6 && face < 0; ++f) {
      const int d = f % 3;
      const bool out = f < 3 ? up[d] < 0 : up[d] > nd[d] - 1;
      if (out && g.face_code[f] != kFacePeriodic) face = f;
    }
    bool up_in = true, dn_in = true;
    for (int d = 0; d < 3; ++d) {
      if (up[d] < 0 && g.face_code[d] == kFacePeriodic) up[d] += nd[d];
      if (up[d] > nd[d] - 1 && g.face_code[d + 3] == kFacePeriodic) up[d] -= nd[d];
      if (dn[d] < 0 && g.face_code[d] == kFacePeriodic) dn[d] += nd[d];
      if (dn[d] > nd[d] - 1 && g.face_code[d + 3] == kFacePeriodic) dn[d] -= nd[d];
      up_in = up_in && up[d] >= 0 && up[d] < nd[d];
      dn_in = dn_in && dn[d] >= 0 && dn[d] < nd[d];
    }
    int up_solid;
    if (up_in) {
      up_solid = state[up[0] + (int64_t)nd[0] * (up[1] + (int64_t)nd[1] * up[2])] < -1000;
    } else {
      int tie = 0;
      up_solid = classify_pos(g, up[0], up[1], up[2], &tie);
      ties += tie;
    }
    const int32_t down = dn_in ? state[dn[0] + (int64_t)nd[0] * (dn[1] + (int64_t)nd[1] * dn[2])] : 0;
    int32_t code;
    if (up_solid) {
      int tie = 0;
      code = wall_code(g, x[0], x[1], x[2], -cv[0], -cv[1], -cv[2], &tie);
      ties += tie;
    } else if (face >= 0) {
      code = g.face_code[face] == -6 ? down : g.face_code[face];
    } else {
      code = down;
    }
    *slot = code == -1001 ? 0 : code;
  }
  if (ties) atomicAdd(counts + 1, (unsigned long long)ties);
}

// node flags: interior fluid (every link a fluid code), face (fluid whose
// negative links all leave the box with exactly the face's code, so the
// streaming kernel can rebuild them from coordinates), boundary (any other
// negative link: wall q codes, null links -- the boundary kernel), solid/null
__global__ void k_flags(const int32_t* __restrict__ state, Box b, FaceCodes fc,
                        uint8_t* __restrict__ flags) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= b.n) return;
  const int64_t n = b.n;
  uint8_t fl = kFlagInterior;
  if (state[k] < 0) {
    fl = kFlagSolid;
  } else {
    const int a = (int)(k % b.nx);
    const int64_t yz = k / b.nx;
    const int bb = (int)(yz % b.ny), c = (int)(yz / b.ny);
    for (int i = 1; i < 27; ++i) {
      const int32_t s = state[(int64_t)i * n + k];
      if (s >= 0) continue;
      const int na = a + CX27[i], nb = bb + CY27[i], nc = c + CZ27[i];
      const bool out[6] = {na < 0 && !b.per[0], nb < 0 && !b.per[1], nc < 0 && !b.per[2],
                           na > b.nx - 1 && !b.per[0], nb > b.ny - 1 && !b.per[1],
                           nc > b.nz - 1 && !b.per[2]};
      int32_t want = 1;
      for (int f = 5; f >= 0; --f)
        if (out[f]) want = fc.code[f];
      if (want == s) {
        if (fl == kFlagInterior) fl = kFlagFace;
      } else {
        fl = kFlagBoundary;
      }
    }
  }
  flags[k] = fl;
}

inline unsigned blocks_for(int64_t n, int t) { return (unsigned)((n + t - 1) / t); }

}  // namespace

hipError_t launch_build_links27(const LinkGeom& g, int32_t* state, unsigned long long* counts,
                                hipStream_t s) {
  k_classify<<<blocks_for(g.box.n, 256), 256, 0, s>>>(g, state, counts);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  k_links<<<blocks_for(g.box.n, 256), 256, 0, s>>>(g, state, counts);
  return hipGetLastError();
}

hipError_t launch_build_links39(const LinkGeom& g, int32_t* state, unsigned long long* counts,
                                hipStream_t s) {
  hipError_t e = hipMemsetAsync(state, 0, (size_t)g.box.n * 39 * sizeof(int32_t), s);
  if (e != hipSuccess) return e;
  k_classify<<<blocks_for(g.box.n, 256), 256, 0, s>>>(g, state, counts);
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  k_links39<<<blocks_for(g.box.n, 256), 256, 0, s>>>(g, state, counts);
  return hipGetLastError();
}

hipError_t launch_build_links39s(const LinkGeom& g, int32_t* state, unsigned long long* counts,
                                 hipStream_t s) {
  hipError_t e = hipMemsetAsync(state, 0, (size_t)g.box.n * 39 * sizeof(int32_t), s);
  if (e != hipSuccess) return e;
  k_classify<<<blocks_for(g.box.n, 256), 256, 0, s>>>(g, state, counts);
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  k_links39s<<<blocks_for(g.box.n, 256), 256, 0, s>>>(g, state, counts);
  return hipGetLastError();
}

// valid nodes of a level of many boxes (flag kFlagInterior on entry); the
// shell keeps its ghost / mirror flags
__global__ void k_flags_boxes(const int32_t* __restrict__ state, int64_t n,
                              uint8_t* __restrict__ flags) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n || flags[k] != kFlagInterior) return;
  uint8_t fl = kFlagInterior;
  if (state[k] < 0) {
    fl = kFlagSolid;
  } else {
    for (int i = 1; i < 27; ++i)
      if (state[(int64_t)i * n + k] < 0) fl = kFlagBoundary;
  }
  flags[k] = fl;
}

hipError_t launch_flags_boxes27(const int32_t* state, int64_t n, uint8_t* flags, hipStream_t s) {
  k_flags_boxes<<<blocks_for(n, 256), 256, 0, s>>>(state, n, flags);
  return hipGetLastError();
}

hipError_t launch_flags27(const int32_t* state, const Box& b, const int32_t face_code[6],
                          uint8_t* flags, hipStream_t s) {
  FaceCodes fc;
  for (int f = 0; f < 6; ++f) fc.code[f] = face_code[f];
  k_flags<<<blocks_for(b.n, 256), 256, 0, s>>>(state, b, fc, flags);
  return hipGetLastError();
}

namespace {

struct BvhBuild {
  const double* tbb;
  std::vector<BvhNode>* nodes;
  std::vector<int32_t>* idx;
  std::vector<double> cen;  // [ntri][3] box centres
  int depth = 0;

  int build(int lo, int hi, int level) {
    if (level > depth) depth = level;
    const int me = (int)nodes->size();
    nodes->push_back(BvhNode{});
    BvhNode nd{};
    for (int d = 0; d < 3; ++d) {
      nd.bb[2 * d] = -1e300;
      nd.bb[2 * d + 1] = 1e300;
    }
    double cmin[3] = {1e300, 1e300, 1e300}, cmax[3] = {-1e300, -1e300, -1e300};
    for (int k = lo; k < hi; ++k) {
      const int t = (*idx)[k];
      for (int d = 0; d < 3; ++d) {
        if (tbb[t * 6 + 2 * d] > nd.bb[2 * d]) nd.bb[2 * d] = tbb[t * 6 + 2 * d];
        if (tbb[t * 6 + 2 * d + 1] < nd.bb[2 * d + 1]) nd.bb[2 * d + 1] = tbb[t * 6 + 2 * d + 1];
        cmin[d] = std::min(cmin[d], cen[t * 3 + d]);
        cmax[d] = std::max(cmax[d], cen[t * 3 + d]);
      }
    }
    if (hi - lo <= 4) {
      nd.a = lo;
      nd.b = -(hi - lo);
      (*nodes)[me] = nd;
      return me;
    }
    int ax = 0;
    for (int d = 1; d < 3; ++d)
      if (cmax[d] - cmin[d] > cmax[ax] - cmin[ax]) ax = d;
    const int mid = (lo + hi) / 2;
    std::nth_element(idx->begin() + lo, idx->begin() + mid, idx->begin() + hi,
                     [&](int32_t p, int32_t q) {
                       const double cp = cen[p * 3 + ax], cq = cen[q * 3 + ax];
                       return cp < cq || (cp == cq && p < q);
                     });
    const int l = build(lo, mid, level + 1);
    const int r = build(mid, hi, level + 1);
    nd.a = l;
    nd.b = r;
    (*nodes)[me] = nd;
    return me;
  }
};

}  // namespace

int build_bvh(const double* tbb, int ntri, std::vector<BvhNode>* nodes,
              std::vector<int32_t>* idx) {
  nodes->clear();
  idx->resize((size_t)ntri);
  if (ntri <= 0) return 0;
  BvhBuild b{tbb, nodes, idx, std::vector<double>((size_t)ntri * 3)};
  for (int t = 0; t < ntri; ++t) {
    (*idx)[t] = t;
    for (int d = 0; d < 3; ++d)
      b.cen[t * 3 + d] = 0.5 * (tbb[t * 6 + 2 * d] + tbb[t * 6 + 2 * d + 1]);
  }
  nodes->reserve((size_t)ntri);
  b.build(0, ntri, 0);
  return b.depth;
}

}  // namespace uclbs
