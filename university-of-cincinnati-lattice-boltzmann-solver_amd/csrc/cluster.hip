// Berger-Rigoutsos clustering of tagged nodes into boxes (host code).
//
// A box of the (coarsened) tag array is shrunk to the bounding box of its
// tags and accepted when the tagged fraction reaches the efficiency.
// Otherwise it is cut, recursively, where its signatures (tag counts per
// plane, one array per axis) say the tags part:
//   1. at a hole, a plane with no tags, the one nearest the box's middle on
//      the longest axis that has one;
//   2. else at the strongest inflection of a signature, between planes i and
//      i + 1 where its discrete Laplacian changes sign, strength
//      |lap(i + 1) - lap(i)|, ties to the cut nearest the middle;
//   3. else through the middle of the longest axis.
#include "cluster.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>

namespace uclbs {
namespace {

int floor_div(int a, int b) { return a >= 0 ? a / b : -((-a + b - 1) / b); }

struct Grid {
  int n[3];
  std::vector<uint8_t> v;
  uint8_t at(int x, int y, int z) const { return v[(size_t)x + (size_t)n[0] * ((size_t)y + (size_t)n[1] * z)]; }
};

struct Cluster {
  const Grid& g;
  double eff;
  std::vector<ClusterBox> boxes;

  // tag counts per plane of box b along each axis
  void signatures(const ClusterBox& b, std::vector<int64_t> sig[3]) const {
    for (int d = 0; d < 3; ++d) sig[d].assign(b.hi[d] - b.lo[d] + 1, 0);
    for (int z = b.lo[2]; z <= b.hi[2]; ++z)
      for (int y = b.lo[1]; y <= b.hi[1]; ++y)
        for (int x = b.lo[0]; x <= b.hi[0]; ++x)
          if (g.at(x, y, z)) {
            sig[0][x - b.lo[0]]++;
            sig[1][y - b.lo[1]]++;
            sig[2][z - b.lo[2]]++;
          }
  }

  void split(ClusterBox b) {
    std::vector<int64_t> sig[3];
    signatures(b, sig);
    // shrink to the tags' bounding box
    int64_t ntag = 0;
    for (int64_t s : sig[0]) ntag += s;
    if (ntag == 0) return;
    for (int d = 0; d < 3; ++d) {
      int a = 0, e = (int)sig[d].size() - 1;
      while (sig[d][a] == 0) ++a;
      while (sig[d][e] == 0) --e;
      sig[d] = std::vector<int64_t>(sig[d].begin() + a, sig[d].begin() + e + 1);
      b.hi[d] = b.lo[d] + e;
      b.lo[d] += a;
    }
    int len[3];
    double vol = 1.0;
    for (int d = 0; d < 3; ++d) {
      len[d] = b.hi[d] - b.lo[d] + 1;
      vol *= len[d];
    }
    if ((double)ntag >= eff * vol || (len[0] == 1 && len[1] == 1 && len[2] == 1)) {
      boxes.push_back(b);
      return;
    }
    int order[3] = {0, 1, 2};
    std::stable_sort(order, order + 3, [&](int p, int q) { return len[p] > len[q]; });
    int axis = -1, cut = 0;  // the left part ends at plane cut (box-local)
    // 1. holes
    for (int k = 0; k < 3 && axis < 0; ++k) {
      const int d = order[k];
      int best = -1;
      for (int i = 1; i + 1 < len[d]; ++i)
        if (sig[d][i] == 0 && (best < 0 || std::abs(2 * i - (len[d] - 1)) < std::abs(2 * best - (len[d] - 1))))
          best = i;
      if (best >= 0) {
        axis = d;
        cut = best - 1;
      }
    }
    // 2. inflections
    if (axis < 0) {
      int64_t strength = 0;
      int centre = 0;
      for (int k = 0; k < 3; ++k) {
        const int d = order[k];
        const std::vector<int64_t>& s = sig[d];
        for (int i = 1; i + 2 < len[d]; ++i) {
          const int64_t l0 = s[i - 1] - 2 * s[i] + s[i + 1];
          const int64_t l1 = s[i] - 2 * s[i + 1] + s[i + 2];
          if (!((l0 < 0 && l1 > 0) || (l0 > 0 && l1 < 0))) continue;
          const int64_t st = std::llabs(l1 - l0);
          const int cen = std::abs(2 * i + 1 - (len[d] - 1));
          if (st > strength || (st == strength && axis >= 0 && cen < centre)) {
            strength = st;
            centre = cen;
            axis = d;
            cut = i;
          }
        }
      }
    }
    // 3. bisection of the longest axis
    if (axis < 0) {
      axis = order[0];
      cut = (len[axis] - 1) / 2;
    }
    ClusterBox left = b, right = b;
    left.hi[axis] = b.lo[axis] + cut;
    right.lo[axis] = b.lo[axis] + cut + 1;
    split(left);
    split(right);
  }
};

// cut [a, e] into pieces of at most max nodes (an even share each where the
// blocking grid allows) whose inner faces lie on multiples of bf
void chop_axis(int a, int e, int max, int bf, std::vector<std::pair<int, int>>& out) {
  out.clear();
  const int len = e - a + 1;
  if (max <= 0 || len <= max) {
    out.push_back({a, e});
    return;
  }
  const int np = (len + max - 1) / max;
  int start = a;
  for (int p = 1; start <= e; ++p) {
    int end;  // first index of the next piece
    if (e - start + 1 <= max) {
      end = e + 1;
    } else {
      const int want = p < np ? a + (int)((int64_t)len * p / np) : start + max;
      end = floor_div(std::min(want, start + max), bf) * bf;
      // start + max >= start + bf, so this multiple of bf lies past start
      if (end <= start) end = floor_div(start + max, bf) * bf;
    }
    out.push_back({start, end - 1});
    start = end;
  }
}

}  // namespace

const char* cluster_tags(const uint8_t* tags, const int32_t lo[3], const int32_t hi[3],
                         double efficiency, int blocking, int buffer, int max_size,
                         std::vector<ClusterBox>& out) {
  out.clear();
  int n[3];
  for (int d = 0; d < 3; ++d) {
    if (hi[d] < lo[d]) return "empty region";
    n[d] = hi[d] - lo[d] + 1;
  }
  if (!(efficiency > 0.0 && efficiency <= 1.0)) return "efficiency outside (0, 1]";
  if (blocking < 1) return "blocking factor below 1";
  if (buffer < 0) return "negative buffer";
  if (max_size < 0 || (max_size > 0 && max_size % blocking != 0))
    return "max_size must be 0 or a multiple of the blocking factor";
  const size_t nn = (size_t)n[0] * n[1] * n[2];
  // grow the tags by the buffer, one axis at a time (a cube)
  std::vector<uint8_t> t(tags, tags + nn), u(nn);
  for (int d = 0; d < 3 && buffer > 0; ++d) {
    const size_t stride = d == 0 ? 1 : (d == 1 ? (size_t)n[0] : (size_t)n[0] * n[1]);
    for (size_t k = 0; k < nn; ++k) {
      const int p = (int)((k / stride) % n[d]);
      uint8_t m = 0;
      for (int o = std::max(0, p - buffer); o <= std::min(n[d] - 1, p + buffer) && !m; ++o)
        m = t[k + (size_t)(o - p) * stride] != 0;
      u[k] = m;
    }
    t.swap(u);
  }
  // coarsen by the blocking factor, on a grid anchored at index 0
  Grid g;
  int clo[3];
  for (int d = 0; d < 3; ++d) {
    clo[d] = floor_div(lo[d], blocking);
    g.n[d] = floor_div(hi[d], blocking) - clo[d] + 1;
  }
  g.v.assign((size_t)g.n[0] * g.n[1] * g.n[2], 0);
  for (int z = 0; z < n[2]; ++z)
    for (int y = 0; y < n[1]; ++y)
      for (int x = 0; x < n[0]; ++x)
        if (t[(size_t)x + (size_t)n[0] * ((size_t)y + (size_t)n[1] * z)]) {
          const int cx = floor_div(lo[0] + x, blocking) - clo[0];
          const int cy = floor_div(lo[1] + y, blocking) - clo[1];
          const int cz = floor_div(lo[2] + z, blocking) - clo[2];
          g.v[(size_t)cx + (size_t)g.n[0] * ((size_t)cy + (size_t)g.n[1] * cz)] = 1;
        }
  Cluster cl{g, efficiency, {}};
  ClusterBox all{{0, 0, 0}, {g.n[0] - 1, g.n[1] - 1, g.n[2] - 1}};
  cl.split(all);
  // refine back, clip, chop
  std::vector<std::pair<int, int>> px, py, pz;
  for (const ClusterBox& c : cl.boxes) {
    ClusterBox f;
    for (int d = 0; d < 3; ++d) {
      f.lo[d] = std::max(lo[d], (c.lo[d] + clo[d]) * blocking);
      f.hi[d] = std::min(hi[d], (c.hi[d] + clo[d]) * blocking + blocking - 1);
    }
    chop_axis(f.lo[0], f.hi[0], max_size, blocking, px);
    chop_axis(f.lo[1], f.hi[1], max_size, blocking, py);
    chop_axis(f.lo[2], f.hi[2], max_size, blocking, pz);
    for (auto& a : pz)
      for (auto& b : py)
        for (auto& e : px)
          out.push_back(ClusterBox{{e.first, b.first, a.first}, {e.second, b.second, a.second}});
  }
  return nullptr;
}

}  // namespace uclbs
