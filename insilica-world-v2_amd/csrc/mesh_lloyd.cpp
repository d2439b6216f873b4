// Jacobi Lloyd iteration toward a spherical centroidal Voronoi tessellation, the mesh builder's one heavy loop.
//
// Each free generator moves to the area-weighted mean of its Voronoi polygon's fan triangles (generator, corner k,
// corner k+1), the corners being the circumcentres of the Delaunay triangles (generator, ring[k], ring[k+1]). All
// generators move from the previous iterate (Jacobi), generators 0 and 11 (the poles) stay pinned, and the loop ends
// when the largest move falls below tol (radians) or after max_iters. Follows the reference's lloyd_scvt
// (src/insilica/grid/icosa.py:194-223) and its small-triangle forms (icosa.py:47-98): chords in a tangent frame at
// the generator, radial parts from a.b - 1 = -|b - a|^2 / 2, and the Van Oosterom-Strackee area.
//
// Built with -ffp-contract=off so every product rounds on its own, as NumPy's element-wise arithmetic does; only
// atan2 comes from the C library. Cells are independent within an iteration, so the OpenMP split changes nothing.
#include <cmath>
#include <cstdint>
#include <vector>

namespace {

struct V3 { double x, y, z; };

inline V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline double norm(V3 a) { return std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }

// Orthonormal tangent pair at unit vector a: e1 = a x (axis least aligned with a), normalised; e2 = a x e1.
inline void frame(V3 a, V3& e1, V3& e2) {
  const double ax = std::fabs(a.x), ay = std::fabs(a.y), az = std::fabs(a.z);
  int k = 0;
  double m = ax;
  if (ay < m) { m = ay; k = 1; }
  if (az < m) { k = 2; }
  const V3 axis = {k == 0 ? 1.0 : 0.0, k == 1 ? 1.0 : 0.0, k == 2 ? 1.0 : 0.0};
  e1 = cross(a, axis);
  const double n = norm(e1);
  e1 = {e1.x / n, e1.y / n, e1.z / n};
  e2 = cross(a, e1);
}

inline V3 circumcentre(V3 a, V3 b, V3 c) {
  V3 e1, e2;
  frame(a, e1, e2);
  const V3 u = sub(b, a), v = sub(c, a);
  const double u1 = dot(u, e1), u2 = dot(u, e2), u3 = -0.5 * dot(u, u);
  const double v1 = dot(v, e1), v2 = dot(v, e2), v3 = -0.5 * dot(v, v);
  const double w1 = u2 * v3 - u3 * v2, w2 = u3 * v1 - u1 * v3, w3 = u1 * v2 - u2 * v1;
  const double t1 = w1 / w3, t2 = w2 / w3;
  V3 p = {a.x + t1 * e1.x + t2 * e2.x, a.y + t1 * e1.y + t2 * e2.y, a.z + t1 * e1.z + t2 * e2.z};
  const double n = norm(p);
  return {p.x / n, p.y / n, p.z / n};
}

inline double tri_area(V3 a, V3 b, V3 c) {
  V3 e1, e2;
  frame(a, e1, e2);
  const V3 u = sub(b, a), v = sub(c, a), w = sub(c, b);
  const double triple = dot(u, e1) * dot(v, e2) - dot(u, e2) * dot(v, e1);
  const double den = 4 - 0.5 * (dot(u, u) + dot(w, w) + dot(v, v));
  return 2 * std::atan2(std::fabs(triple), den);
}

}  // namespace

extern "C" int isw_mesh_lloyd(int N, double* pos, const int32_t* ring, const int32_t* deg, int max_iters, double tol,
                              int* iters_out, double* moved_out) {
  std::vector<V3> cur(N), nxt(N);
  for (int i = 0; i < N; i++) cur[i] = {pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]};
  int it = 0;
  double moved = INFINITY;
  for (it = 1; it <= max_iters; it++) {
    double mv = 0;
#pragma omp parallel for reduction(max : mv) schedule(static)
    for (int i = 0; i < N; i++) {
      const V3 a = cur[i];
      const int m = deg[i];
      V3 C[6];
      for (int k = 0; k < m; k++) C[k] = circumcentre(a, cur[ring[6 * i + k]], cur[ring[6 * i + (k + 1) % m]]);
      V3 s = {0, 0, 0};
      for (int k = 0; k < m; k++) {
        const V3 c0 = C[k], c1 = C[(k + 1) % m];
        const double ar = tri_area(a, c0, c1);
        s = {s.x + ar * (a.x + c0.x + c1.x), s.y + ar * (a.y + c0.y + c1.y), s.z + ar * (a.z + c0.z + c1.z)};
      }
      const double n = norm(s);
      const V3 p = {s.x / n, s.y / n, s.z / n};
      if (i == 0 || i == 11) {
        nxt[i] = a;
      } else {
        const double d = norm(sub(p, a));
        if (d > mv) mv = d;
        nxt[i] = p;
      }
    }
    cur.swap(nxt);
    moved = mv;
    if (moved < tol) break;
  }
  if (it > max_iters) it = max_iters;
  for (int i = 0; i < N; i++) { pos[3 * i] = cur[i].x; pos[3 * i + 1] = cur[i].y; pos[3 * i + 2] = cur[i].z; }
  *iters_out = it;
  *moved_out = moved;
  return 0;
}
