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

namespace {

// point of a wedge face from its parameters: corners p0 -> p1 along the face,
// z0 -> z1 across the span; s, t exactly 0 or 1 give the corner lines exactly,
// so faces that share an edge share its vertices bit for bit
V3 face_point(V3 p0, V3 p1, double s, double z0, double z1, double t) {
  V3 p = add(scl(p0, 1.0 - s), scl(p1, s));
  p.z = z0 * (1.0 - t) + z1 * t;
  return p;
}

}  // namespace

extern "C" int uclbs_geom_wedge(double apex_x, double apex_y, double length,
                                double half_angle_deg, double z0, double z1, int n_len,
                                int n_span, uint64_t seed, double* tri, double* nrm,
                                int64_t capacity, int64_t* ntri) {
  if (!(length > 0.0) || !(half_angle_deg > 0.0 && half_angle_deg < 60.0) || !(z1 > z0) ||
      n_len < 1 || n_span < 1 || n_len > 4096 || n_span > 4096 || !ntri)
    return UCLBS_E_ARG;
  const int64_t count = 6 * (int64_t)n_len * n_span + 6 * (int64_t)n_len;
  *ntri = count;
  if (!tri || !nrm) return UCLBS_OK;  // size query
  if (capacity < count) return UCLBS_E_ARG;
  const double h = length * std::tan(half_angle_deg * 3.14159265358979323846 / 180.0);
  const V3 A{apex_x, apex_y, 0.0}, B{apex_x + length, apex_y + h, 0.0},
      C{apex_x + length, apex_y - h, 0.0};
  const V3 cen{(A.x + B.x + C.x) / 3.0, (A.y + B.y + C.y) / 3.0, 0.5 * (z0 + z1)};
  int64_t w = 0;
  auto emit = [&](V3 a, V3 b, V3 d) {
    V3 n = cross(sub(b, a), sub(d, a));
    if (dot(n, sub(a, cen)) < 0.0) {  // outward orientation (the wedge is convex)
      std::swap(b, d);
      n = scl(n, -1.0);
    }
    n = unit(n);
    const V3 vs[3] = {a, b, d};
    for (int k = 0; k < 3; ++k) {
      tri[w * 9 + k * 3 + 0] = vs[k].x;
      tri[w * 9 + k * 3 + 1] = vs[k].y;
      tri[w * 9 + k * 3 + 2] = vs[k].z;
    }
    nrm[w * 3 + 0] = n.x;
    nrm[w * 3 + 1] = n.y;
    nrm[w * 3 + 2] = n.z;
    ++w;
  };
  // the three side faces (upper A->B, base B->C, lower C->A), each an
  // n_len x n_span grid whose interior vertices are moved inside the face
  // plane by up to a quarter cell (Philox uniforms from the seed): the surface
  // is the exact prism whatever the seed, the triangle edges are not
  const V3 corner[4] = {A, B, C, A};
  std::vector<V3> grid((size_t)(n_len + 1) * (n_span + 1));
  for (int f = 0; f < 3; ++f) {
    for (int j = 0; j <= n_span; ++j)
      for (int i = 0; i <= n_len; ++i) {
        double s = (double)i / n_len, t = (double)j / n_span;
        if (i > 0 && i < n_len && j > 0 && j < n_span) {
          const uint64_t id = ((uint64_t)f * 4097 + (uint64_t)j) * 4097 + (uint64_t)i;
          s += (uclbs::philox_uniform(id, 0, seed) - 0.5) * 0.5 / n_len;
          t += (uclbs::philox_uniform(id, 1, seed) - 0.5) * 0.5 / n_span;
        }
        if (i == n_len) s = 1.0;
        if (j == n_span) t = 1.0;
        grid[(size_t)j * (n_len + 1) + i] = face_point(corner[f], corner[f + 1], s, z0, z1, t);
      }
    for (int j = 0; j < n_span; ++j)
      for (int i = 0; i < n_len; ++i) {
        const V3 v00 = grid[(size_t)j * (n_len + 1) + i], v10 = grid[(size_t)j * (n_len + 1) + i + 1];
        const V3 v01 = grid[(size_t)(j + 1) * (n_len + 1) + i],
                 v11 = grid[(size_t)(j + 1) * (n_len + 1) + i + 1];
        emit(v00, v10, v11);
        emit(v00, v11, v01);
      }
  }
  // end caps: fans from the cap centroid over the boundary points the side
  // faces put on the edges at z0 and z1 (no T-junctions)
  for (int e = 0; e < 2; ++e) {
    const double t = (double)e, ze = e ? z1 : z0;
    const V3 cc{cen.x, cen.y, ze};
    for (int f = 0; f < 3; ++f)
      for (int i = 0; i < n_len; ++i) {
        const double s0 = (double)i / n_len, s1 = i + 1 == n_len ? 1.0 : (double)(i + 1) / n_len;
        emit(cc, face_point(corner[f], corner[f + 1], s0, z0, z1, t),
             face_point(corner[f], corner[f + 1], s1, z0, z1, t));
      }
  }
  return UCLBS_OK;
}
