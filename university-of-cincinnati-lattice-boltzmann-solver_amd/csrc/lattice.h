// Lattice constant tables shared by host and device code.
//
// D3Q27 order is linking/modules/linkwise.f95:30-37 (cx_27, cy_27, cz_27, opp_27),
// 0-based.  The same order is what directions_and_weights.f95:336-343 assigns to
// cx/cy/cz when dir = 27, so the link codes of a reference `state` array for a
// D3Q27 run index these tables directly (component q+1 <-> index q).
// D3Q39 order is directions_and_weights.f95:436-444 with opp :498-501.
#pragma once
#include <cstdint>

namespace uclbs {

constexpr int Q27 = 27;
constexpr int CX27[27] = {0, 1, 0, 0, -1, 0, 0, 1, 0, -1, 0, 1, 1, -1, -1, 1, 0, -1, 0,
                          1, 1, 1, 1, -1, -1, -1, -1};
constexpr int CY27[27] = {0, 0, 1, 0, 0, -1, 0, 1, 1, 1, 1, 0, 0, 0, 0, -1, -1, -1, -1,
                          1, 1, -1, -1, 1, 1, -1, -1};
constexpr int CZ27[27] = {0, 0, 0, 1, 0, 0, -1, 0, -1, 0, 1, 1, -1, -1, 1, 0, -1, 0, 1,
                          1, -1, 1, -1, 1, -1, 1, -1};
constexpr int OPP27[27] = {0, 4, 5, 6, 1, 2, 3, 17, 18, 15, 16, 13, 14, 11, 12,
                           9, 10, 7, 8, 26, 25, 24, 23, 22, 21, 20, 19};
// w = w1(cx) w1(cy) w1(cz) with w1(0) = 2/3, w1(+-1) = 1/6
constexpr double W27_0 = 8.0 / 27.0, W27_1 = 2.0 / 27.0, W27_2 = 1.0 / 54.0,
                 W27_3 = 1.0 / 216.0;
constexpr double W27[27] = {W27_0, W27_1, W27_1, W27_1, W27_1, W27_1, W27_1, W27_2, W27_2,
                            W27_2, W27_2, W27_2, W27_2, W27_2, W27_2, W27_2, W27_2, W27_2,
                            W27_2, W27_3, W27_3, W27_3, W27_3, W27_3, W27_3, W27_3, W27_3};

constexpr int Q39 = 39;
constexpr int CX39[39] = {0, 1, 0, 0, -1, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1, 2, 0, 0, -2, 0,
                          0, 2, 2, 0, -2, -2, 2, 2, -2, -2, 0, 0, 0, 3, 0, 0, -3, 0, 0};
constexpr int CY39[39] = {0, 0, 1, 0, 0, -1, 0, 1, 1, -1, -1, 1, 1, -1, -1, 0, 2, 0, 0, -2,
                          0, 2, 0, 2, 2, -2, -2, 0, 0, 0, 2, -2, -2, 0, 3, 0, 0, -3, 0};
constexpr int CZ39[39] = {0, 0, 0, 1, 0, 0, -1, 1, 1, 1, 1, -1, -1, -1, -1, 0, 0, 2, 0, 0,
                          -2, 0, 2, 2, 0, 0, 0, -2, -2, 2, -2, -2, 2, 0, 0, 3, 0, 0, -3};
constexpr int OPP39[39] = {0, 4, 5, 6, 1, 2, 3, 13, 14, 11, 12, 9, 10, 7, 8, 18, 19, 20, 15, 16,
                           17, 25, 28, 31, 26, 21, 24, 29, 22, 27, 32, 23, 30, 36, 37, 38, 33,
                           34, 35};

// Stencil kind of the shifted-lattice wall derivatives of link i along axis a
// (derivatives.f95:816-1235): 1 for the case's dudx_ibb with winds = .false.,
// -1 with winds = .true., 0 for the local difference.  The sign of the
// unshifted component (the rest direction: + along x), except for (2,0,-2),
// (-2,0,2), (0,2,-2), (0,-2,2), whose cases carry the reversed kinds.
constexpr signed char WALLKIND39[39][3] = {
    {1, 0, 0},   {1, 0, 0},  {0, 1, 0},   {0, 0, 1},   {-1, 0, 0},  {0, -1, 0}, {0, 0, -1},
    {1, 1, 1},   {-1, 1, 1}, {-1, -1, 1}, {1, -1, 1},  {1, 1, -1},  {-1, 1, -1}, {-1, -1, -1},
    {1, -1, -1}, {1, 0, 0},  {0, 1, 0},   {0, 0, 1},   {-1, 0, 0},  {0, -1, 0}, {0, 0, -1},
    {1, 1, 0},   {1, 0, 1},  {0, 1, 1},   {-1, 1, 0},  {-1, -1, 0}, {1, -1, 0}, {-1, 0, 1},
    {-1, 0, -1}, {1, 0, -1}, {0, -1, 1},  {0, -1, -1}, {0, 1, -1},  {1, 0, 0},  {0, 1, 0},
    {0, 0, 1},   {-1, 0, 0}, {0, -1, 0},  {0, 0, -1}};

// kMirror39.m[d][k]: the direction whose velocity is c_k with component d negated
// (the specular image at a wall normal to axis d).  Built from the un-shifted
// velocities; for d = y, z it holds on the shifted lattice as well, whose
// shift is along x.
struct Mirror39 {
  int m[3][39];
};
constexpr Mirror39 make_mirror39() {
  Mirror39 r{};
  for (int d = 0; d < 3; ++d)
    for (int k = 0; k < 39; ++k) {
      const int c[3] = {CX39[k], CY39[k], CZ39[k]};
      for (int j = 0; j < 39; ++j) {
        const int e[3] = {CX39[j], CY39[j], CZ39[j]};
        bool same = true;
        for (int a = 0; a < 3; ++a) same = same && (e[a] == (a == d ? -c[a] : c[a]));
        if (same) r.m[d][k] = j;
      }
    }
  return r;
}
constexpr Mirror39 kMirror39 = make_mirror39();

}  // namespace uclbs
