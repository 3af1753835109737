// Triangulated geometry inputs (host code).
//
// The reference imports a binary STL into tri_vertices(3,3,ntri) and
// normal(3,ntri) (geometry/geom_import.f95:48-59,133; geometry/modules/
// geom_data.f95:9-12).  The benchmark geometries are generated instead, in the
// same triangle-soup layout: tri[t*9 + v*3 + axis], nrm[t*3 + axis].
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <vector>

#include "../../include/uclbs_hip.h"
#include "philox.h"

namespace {

struct V3 {
  double x, y, z;
};
V3 add(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
V3 scl(V3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
V3 unit(V3 a) { return scl(a, 1.0 / std::sqrt(dot(a, a))); }

struct Tri {
  V3 a, b, c;
};

void subdivide(std::vector<Tri>& t) {
  std::vector<Tri> o;
  o.reserve(t.size() * 4);
  for (const Tri& s : t) {
    // a+b and b+a round identically, so shared edges get identical midpoints
    V3 ab = unit(scl(add(s.a, s.b), 0.5)), bc = unit(scl(add(s.b, s.c), 0.5)),
       ca = unit(scl(add(s.c, s.a), 0.5));
    o.push_back({s.a, ab, ca});
    o.push_back({ab, s.b, bc});
    o.push_back({ca, bc, s.c});
    o.push_back({ab, bc, ca});
  }
  t.swap(o);
}

// rotate v by unit quaternion (w, x, y, z)
V3 qrot(const double q[4], V3 v) {
  V3 u{q[1], q[2], q[3]};
  V3 t = scl(cross(u, v), 2.0);
  return add(add(v, scl(t, q[0])), cross(u, t));
}

}  // namespace

extern "C" int uclbs_geom_icosphere(int level, double radius, const double center[3],
                                    uint64_t seed, double* tri, double* nrm, int64_t capacity,
                                    int64_t* ntri) {
  if (level < 0 || level > 8 || !(radius > 0.0) || !center || !ntri) return UCLBS_E_ARG;
  int64_t count = 20;
  for (int l = 0; l < level; ++l) count *= 4;
  *ntri = count;
  if (!tri || !nrm) return UCLBS_OK;  // size query
  if (capacity < count) return UCLBS_E_ARG;
  const double p = (1.0 + std::sqrt(5.0)) / 2.0;
  const V3 v[12] = {{-1, p, 0}, {1, p, 0}, {-1, -p, 0}, {1, -p, 0}, {0, -1, p}, {0, 1, p},
                    {0, -1, -p}, {0, 1, -p}, {p, 0, -1}, {p, 0, 1}, {-p, 0, -1}, {-p, 0, 1}};
  const int f[20][3] = {{0, 11, 5}, {0, 5, 1}, {0, 1, 7}, {0, 7, 10}, {0, 10, 11},
                        {1, 5, 9}, {5, 11, 4}, {11, 10, 2}, {10, 7, 6}, {7, 1, 8},
                        {3, 9, 4}, {3, 4, 2}, {3, 2, 6}, {3, 6, 8}, {3, 8, 9},
                        {4, 9, 5}, {2, 4, 11}, {6, 2, 10}, {8, 6, 7}, {9, 8, 1}};
  std::vector<Tri> t;
  for (auto& k : f) t.push_back({unit(v[k[0]]), unit(v[k[1]]), unit(v[k[2]])});
  for (int l = 0; l < level; ++l) subdivide(t);
  // uniform random rotation (Shoemake) from three Philox uniforms
  const double u1 = uclbs::philox_uniform(0, 0, seed), u2 = uclbs::philox_uniform(1, 0, seed),
               u3 = uclbs::philox_uniform(2, 0, seed);
  const double tp = 6.283185307179586476925286766559;
  const double q[4] = {std::sqrt(u1) * std::cos(tp * u3), std::sqrt(1.0 - u1) * std::sin(tp * u2),
                       std::sqrt(1.0 - u1) * std::cos(tp * u2), std::sqrt(u1) * std::sin(tp * u3)};
  const V3 c{center[0], center[1], center[2]};
  for (int64_t i = 0; i < count; ++i) {
    V3 a = add(c, scl(qrot(q, t[i].a), radius));
    V3 b = add(c, scl(qrot(q, t[i].b), radius));
    V3 d = add(c, scl(qrot(q, t[i].c), radius));
    V3 n = cross(sub(b, a), sub(d, a));
    if (dot(n, sub(a, c)) < 0.0) {  // outward orientation
      std::swap(b, d);
      n = scl(n, -1.0);
    }
    n = unit(n);
    const V3 vs[3] = {a, b, d};
    for (int k = 0; k < 3; ++k) {
      tri[i * 9 + k * 3 + 0] = vs[k].x;
      tri[i * 9 + k * 3 + 1] = vs[k].y;
      tri[i * 9 + k * 3 + 2] = vs[k].z;
    }
    nrm[i * 3 + 0] = n.x;
    nrm[i * 3 + 1] = n.y;
    nrm[i * 3 + 2] = n.z;
  }
  return UCLBS_OK;
}
