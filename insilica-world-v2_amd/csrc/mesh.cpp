// The planet's mesh on the host: the icosahedral Voronoi grid smoothed toward a spherical centroidal Voronoi
// tessellation, and the TRiSK tables the HIP kernels read. libinsilica_mesh.so; Python (insilica/grid) allocates
// every output and calls the stages in order.
//
// A drop-in mesh must number its cells, edges and vertices as the reference's generator does and agree with its
// geometry to the bit (src/insilica/grid/icosa.py, grid/trisk.py), so these stages keep its rules and its
// floating-point forms: the icosahedron with a vertex at each pole, subdivided by edge midpoints numbered as the
// faces are walked; neighbour rings sorted counter-clockwise by angle from local east; Jacobi Lloyd with the poles
// pinned; each Voronoi corner (a Delaunay triangle's circumcentre) computed once and shared by its three cells (the
// first cell in cell order wins); edges listed by lower cell then slot; vertices numbered by first appearance over
// (cell, slot); the tangential weights walked counter-clockwise from each edge around both its cells; the 3-D Hilbert
// key of Skilling (2004) on 21-bit quantised positions. The small-triangle forms (chords in a tangent frame,
// a.b - 1 = -|b - a|^2 / 2, Van Oosterom and Strackee's area) are computed here; the element-wise arcsin and arctan2
// that turn chords and (triple product, denominator) pairs into angles run in NumPy, whose vectorised versions the
// reference's arrays come from (the C library's round differently in a few per cent of cases). Lloyd's loop, which
// needs arctan2 every iteration, uses the C library's (the iteration counts and fixtures agree).
//
// Built with g++ -O2 -ffp-contract=off (every product rounds on its own, as NumPy's element-wise arithmetic does) and
// OpenMP over cells where cells are independent.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <unordered_map>
#include <vector>

namespace {

struct V3 {
  double x, y, z;
};
inline V3 add(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 scale(V3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
inline V3 divide(V3 a, double s) { return {a.x / s, a.y / s, a.z / s}; }
inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline double norm(V3 a) { return std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }
inline V3 get3(const double* p, int64_t i) { return {p[3 * i], p[3 * i + 1], p[3 * i + 2]}; }
inline void put3(double* p, int64_t i, V3 v) { p[3 * i] = v.x; p[3 * i + 1] = v.y; p[3 * i + 2] = v.z; }

// Orthonormal tangent pair at unit vector a: e1 = a x (the coordinate axis least aligned with a, the first of equals),
// normalised; e2 = a x e1.
inline void frame(V3 a, V3& e1, V3& e2) {
  const double ax = std::fabs(a.x), ay = std::fabs(a.y), az = std::fabs(a.z);
  int k = 0;
  double m = ax;
  if (ay < m) { m = ay; k = 1; }
  if (az < m) k = 2;
  const V3 axis = {k == 0 ? 1.0 : 0.0, k == 1 ? 1.0 : 0.0, k == 2 ? 1.0 : 0.0};
  e1 = cross(a, axis);
  e1 = divide(e1, norm(e1));
  e2 = cross(a, e1);
}

// The unit circumcentre of the spherical triangle (a, b, c): the plane normal built from the chords in a's frame.
inline V3 circumcentre(V3 a, V3 b, V3 c) {
  V3 e1, e2;
  frame(a, e1, e2);
  const V3 u = sub(b, a), v = sub(c, a);
  const double u1 = dot(u, e1), u2 = dot(u, e2), u3 = -0.5 * dot(u, u);
  const double v1 = dot(v, e1), v2 = dot(v, e2), v3 = -0.5 * dot(v, v);
  const double w1 = u2 * v3 - u3 * v2, w2 = u3 * v1 - u1 * v3, w3 = u1 * v2 - u2 * v1;
  const double t1 = w1 / w3, t2 = w2 / w3;
  const V3 p = add(add(a, scale(e1, t1)), scale(e2, t2));
  return divide(p, norm(p));
}

// The spherical triangle (a, b, c)'s area is 2 atan2(|t|, d): t = a.(b x c) from the chords in a's frame, d = 1 + a.b
// + b.c + c.a written as 4 - (|b - a|^2 + |c - b|^2 + |a - c|^2) / 2.
inline void tri_parts(V3 a, V3 b, V3 c, double& t, double& d) {
  V3 e1, e2;
  frame(a, e1, e2);
  const V3 u = sub(b, a), v = sub(c, a), w = sub(c, b);
  t = dot(u, e1) * dot(v, e2) - dot(u, e2) * dot(v, e1);
  d = 4 - 0.5 * (dot(u, u) + dot(w, w) + dot(v, v));
}
inline double tri_area_libm(V3 a, V3 b, V3 c) {
  double t, d;
  tri_parts(a, b, c, t, d);
  return 2 * std::atan2(std::fabs(t), d);
}

// Local east and north at unit vector p; east is (0, 1, 0) at the exact poles.
inline void east_north(V3 p, V3& east, V3& north) {
  double ex = -p.y, ey = p.x;
  double er = std::sqrt(ex * ex + ey * ey);
  if (er < 1e-12) { ex = 0.0; ey = 1.0; er = 1.0; }
  east = {ex / er, ey / er, 0.0};
  north = cross(p, east);
}

// Each cell's neighbours sorted counter-clockwise by angle from local east (stable), pads (-1) last.
void ring_sort(int64_t N, const double* pos, const int32_t* adj, const uint8_t* deg, int32_t* ring) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < N; i++) {
    V3 east, north;
    east_north(get3(pos, i), east, north);
    const int m = deg[i];
    double ang[6];
    int idx[6];
    for (int k = 0; k < m; k++) {
      const V3 q = get3(pos, adj[6 * i + k]);
      ang[k] = std::atan2(dot(q, north), dot(q, east));
      idx[k] = k;
    }
    std::stable_sort(idx, idx + m, [&](int a, int b) { return ang[a] < ang[b]; });
    for (int k = 0; k < 6; k++) ring[6 * i + k] = k < m ? adj[6 * i + idx[k]] : -1;
  }
}

inline uint64_t edge_key(int64_t a, int64_t b) { return a < b ? (uint64_t)a << 32 | (uint64_t)b : (uint64_t)b << 32 | (uint64_t)a; }

}  // namespace

extern "C" {

// N, E, NV and the tangential weights' count at level n.
int isw_mesh_sizes(int n, int64_t* N, int64_t* E, int64_t* NV, int64_t* n_eoe) {
  if (n < 0 || n > 10) return -1;
  const int64_t f = (int64_t)1 << (2 * n);
  *N = 10 * f + 2;
  *E = 30 * f;
  *NV = 20 * f;
  // each edge walks its two cells' other edges: sum over edges of (deg(eI) - 1) + (deg(eJ) - 1) = sum over cells of
  // deg (deg - 1), with 12 pentagons and N - 12 hexagons
  *n_eoe = 12 * 5 * 4 + (*N - 12) * 6 * 5;
  return 0;
}

// Stage 1: the icosphere at level n from the 12 base vertices base (pole, upper ring, lower ring, pole), subdivided
// by edge midpoints (each new point normalised by pow(x x + y y + z z, 0.5), numbered as the faces are walked), its
// adjacency (each cell's neighbours, ascending, pads -1) and degrees. pos (N, 3), adj (N, 6), deg (N).
int isw_mesh_icosphere(int n, const double* base, double* pos, int32_t* adj, uint8_t* deg) {
  int64_t N, E, NV, ne;
  if (isw_mesh_sizes(n, &N, &E, &NV, &ne)) return -1;
  std::vector<V3> P;
  P.reserve(N);
  auto put = [&](double x, double y, double z) {
    const double r = std::pow(x * x + y * y + z * z, 0.5);
    P.push_back({x / r, y / r, z / r});
    return (int)P.size() - 1;
  };
  for (int k = 0; k < 12; k++) put(base[3 * k], base[3 * k + 1], base[3 * k + 2]);
  std::vector<std::array<int, 3>> faces;
  for (int k = 0; k < 5; k++) {
    const int a = 1 + k, b = 1 + (k + 1) % 5, c = 6 + k, d = 6 + (k + 1) % 5;
    faces.push_back({0, a, b});
    faces.push_back({a, c, b});
    faces.push_back({b, c, d});
    faces.push_back({11, d, c});
  }
  for (int level = 0; level < n; level++) {
    std::unordered_map<uint64_t, int> mid;
    mid.reserve(faces.size() * 2);
    auto midpoint = [&](int i, int j) {
      const uint64_t key = edge_key(i, j);
      auto it = mid.find(key);
      if (it != mid.end()) return it->second;
      const V3 p = P[i], q = P[j];
      const int m = put(p.x + q.x, p.y + q.y, p.z + q.z);
      mid.emplace(key, m);
      return m;
    };
    std::vector<std::array<int, 3>> sub_faces;
    sub_faces.reserve(faces.size() * 4);
    for (const auto& f : faces) {
      const int ab = midpoint(f[0], f[1]), bc = midpoint(f[1], f[2]), ca = midpoint(f[2], f[0]);
      sub_faces.push_back({f[0], ab, ca});
      sub_faces.push_back({f[1], bc, ab});
      sub_faces.push_back({f[2], ca, bc});
      sub_faces.push_back({ab, bc, ca});
    }
    faces.swap(sub_faces);
  }
  if ((int64_t)P.size() != N) return -2;
  for (int64_t i = 0; i < N; i++) put3(pos, i, P[i]);
  // neighbours: every face's three directed pairs, then each cell's distinct neighbours ascending
  std::vector<std::vector<int32_t>> nb(N);
  for (const auto& f : faces)
    for (int a = 0; a < 3; a++)
      for (int b = 0; b < 3; b++)
        if (a != b) nb[f[a]].push_back(f[b]);
  for (int64_t i = 0; i < N; i++) {
    auto& v = nb[i];
    std::sort(v.begin(), v.end());
    v.erase(std::unique(v.begin(), v.end()), v.end());
    if (v.size() < 5 || v.size() > 6) return -3;
    deg[i] = (uint8_t)v.size();
    for (int k = 0; k < 6; k++) adj[6 * i + k] = k < (int)v.size() ? v[k] : -1;
  }
  return 0;
}

// Neighbour rings sorted counter-clockwise from local east (pads -1 last).
int isw_mesh_ring(int64_t N, const double* pos, const int32_t* adj, const uint8_t* deg, int32_t* ring) {
  ring_sort(N, pos, adj, deg, ring);
  return 0;
}

// Jacobi Lloyd toward the SCVT: each free generator to the area-weighted mean of its Voronoi polygon's fan triangles
// (generator, corner k, corner k + 1), all from the previous iterate; generators 0 and 11 (the poles) pinned; until
// the largest move (radians) < tol or max_iters. ring (N, 6) as isw_mesh_ring's. pos updated in place.
int isw_mesh_lloyd(int64_t N, double* pos, const int32_t* ring, const uint8_t* deg, int max_iters, double tol,
                   int* iters_out, double* moved_out) {
  std::vector<V3> cur(N), nxt(N);
  for (int64_t i = 0; i < N; i++) cur[i] = get3(pos, i);
  int it = 0;
  double moved = INFINITY;
  for (it = 1; it <= max_iters; it++) {
    double mv = 0;
#pragma omp parallel for reduction(max : mv) schedule(static)
    for (int64_t i = 0; i < N; i++) {
      const V3 a = cur[i];
      const int m = deg[i];
      V3 C[6];
      for (int k = 0; k < m; k++) C[k] = circumcentre(a, cur[ring[6 * i + k]], cur[ring[6 * i + (k + 1) % m]]);
      V3 s = {0, 0, 0};
      for (int k = 0; k < m; k++) {
        const V3 c0 = C[k], c1 = C[(k + 1) % m];
        s = add(s, scale(add(add(a, c0), c1), tri_area_libm(a, c0, c1)));
      }
      const V3 p = divide(s, norm(s));
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
  for (int64_t i = 0; i < N; i++) put3(pos, i, cur[i]);
  *iters_out = it;
  *moved_out = moved;
  return 0;
}

// Stage 2: the Voronoi cells of the smoothed generators, in the order of the cells given: east and north (N, 3); the
// corners (N, 6, 3) (circumcentres of (cell, ring k, ring k + 1), each computed once by the first cell in this order
// that meets it and copied to the other two; 0 in pad slots); the chords |x_ring(k) - x_i| and |corner k - corner
// k - 1| (N, 6) for the distances and edge lengths; the fan triangles' (triple, denominator) pairs (N, 6, 2) for the
// area. ring (N, 6) sorted, pads -1.
int isw_mesh_cells(int64_t N, const double* pos, const int32_t* ring, const uint8_t* deg, double* east,
                   double* north, double* corners, double* chord_d, double* chord_e, double* area_td) {
  std::memset(corners, 0, sizeof(double) * 18 * N);
  // corners, computed by their first cell
  std::unordered_map<uint64_t, int64_t> owner;      // sorted triple -> flat (cell, slot) of the first copy
  owner.reserve(2 * N);
  auto tri_key = [&](int64_t a, int64_t b, int64_t c) {
    int64_t s[3] = {a, b, c};
    std::sort(s, s + 3);
    return (uint64_t)((s[0] * N + s[1]) * N + s[2]);
  };
  std::vector<int64_t> src(6 * N, -1);
  for (int64_t i = 0; i < N; i++) {
    const int m = deg[i];
    for (int k = 0; k < m; k++) {
      const uint64_t key = tri_key(i, ring[6 * i + k], ring[6 * i + (k + 1) % m]);
      auto it = owner.find(key);
      if (it == owner.end()) {
        owner.emplace(key, 6 * i + k);
        src[6 * i + k] = 6 * i + k;
      } else {
        src[6 * i + k] = it->second;
      }
    }
  }
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < N; i++) {
    const int m = deg[i];
    const V3 p = get3(pos, i);
    V3 e, no;
    east_north(p, e, no);
    put3(east, i, e);
    put3(north, i, no);
    for (int k = 0; k < m; k++)
      if (src[6 * i + k] == 6 * i + k)
        put3(corners, 6 * i + k, circumcentre(p, get3(pos, ring[6 * i + k]), get3(pos, ring[6 * i + (k + 1) % m])));
  }
  for (int64_t f = 0; f < 6 * N; f++)
    if (src[f] >= 0 && src[f] != f) put3(corners, f, get3(corners, src[f]));
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < N; i++) {
    const int m = deg[i];
    const V3 p = get3(pos, i);
    for (int k = 0; k < 6; k++) {
      double* td = area_td + 2 * (6 * i + k);
      if (k >= m) {
        chord_d[6 * i + k] = chord_e[6 * i + k] = 0.0;
        td[0] = td[1] = 0.0;
        continue;
      }
      const V3 ck = get3(corners, 6 * i + k), cp = get3(corners, 6 * i + (k + m - 1) % m);
      chord_d[6 * i + k] = norm(sub(get3(pos, ring[6 * i + k]), p));
      chord_e[6 * i + k] = norm(sub(ck, cp));
      tri_parts(p, ck, get3(corners, 6 * i + (k + 1) % m), td[0], td[1]);
    }
  }
  return 0;
}

// The edges (i < j) listed by lower cell i then i's slot, with each side's slot.
int isw_mesh_edges(int64_t N, const int32_t* ring, const uint8_t* deg, int32_t* eI, int32_t* eJ, uint8_t* eKi,
                   uint8_t* eKj) {
  int64_t e = 0;
  for (int64_t i = 0; i < N; i++)
    for (int k = 0; k < deg[i]; k++) {
      const int64_t j = ring[6 * i + k];
      if (j <= i) continue;
      int kj = 0;
      while (kj < 6 && ring[6 * j + kj] != i) kj++;
      if (kj == 6) return -1;
      eI[e] = (int32_t)i;
      eJ[e] = (int32_t)j;
      eKi[e] = (uint8_t)k;
      eKj[e] = (uint8_t)kj;
      e++;
    }
  return 0;
}

// The cells' order along a 3-D Hilbert curve through their unit positions quantised to `bits` bits per axis
// (Skilling 2004: the axes to the transposed Hilbert index, then its bits interleaved into one key; stable order).
int isw_mesh_hilbert(int64_t N, const double* pos, int bits, int32_t* perm) {
  const uint64_t top = ((uint64_t)1 << bits) - 1;
  std::vector<uint64_t> key(N);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < N; i++) {
    uint64_t X[3];
    for (int a = 0; a < 3; a++) {
      const double q = (pos[3 * i + a] + 1) * 0.5 * (double)top;
      uint64_t v = q <= 0 ? 0 : (uint64_t)q;          // (truncation, as a float-to-unsigned cast)
      X[a] = v > top ? top : v;
    }
    for (uint64_t Q = (uint64_t)1 << (bits - 1); Q > 1; Q >>= 1) {
      const uint64_t low = Q - 1;
      for (int a = 0; a < 3; a++) {
        if (X[a] & Q) {
          X[0] ^= low;
        } else {
          const uint64_t t = (X[0] ^ X[a]) & low;
          X[0] ^= t;
          X[a] ^= t;
        }
      }
    }
    X[1] ^= X[0];
    X[2] ^= X[1];
    uint64_t t = 0;
    for (uint64_t Q = (uint64_t)1 << (bits - 1); Q > 1; Q >>= 1)
      if (X[2] & Q) t ^= Q - 1;
    for (int a = 0; a < 3; a++) X[a] ^= t;
    uint64_t k = 0;
    for (int b = bits - 1; b >= 0; b--)
      for (int a = 0; a < 3; a++) k = (k << 1) | ((X[a] >> b) & 1);
    key[i] = k;
  }
  std::vector<int32_t> idx(N);
  std::iota(idx.begin(), idx.end(), 0);
  std::stable_sort(idx.begin(), idx.end(), [&](int32_t a, int32_t b) { return key[a] < key[b]; });
  std::memcpy(perm, idx.data(), sizeof(int32_t) * N);
  return 0;
}

// Stage 3, the TRiSK tables in the (final) cell order: vertices numbered by first appearance over (cell, slot),
// their positions vPos (NV, 3) (the shared corners); cornerV (N, 6); cellEdge, cellSign (N, 6); ev1, ev2 (E) (the
// corners before and at eI's slot: the tangent runs ev1 -> ev2); xe (E, 3) where the line between the two cells
// crosses the edge, ne (E, 3) the unit normal eI -> eJ; xm (E, 3) the Voronoi edge's midpoint; the kites' triangle
// parts kite_td (N, 6, 2, 2) ((x_i, xe of edge k, vertex k) and (x_i, vertex k, xe of edge k + 1)); circ1, circ2 (E)
// the edge's circulation sign about ev1 and ev2; cellEdgeAsc (N, 6); vEdge, vEdgeSign (NV, 3); vCell, vSlot (NV, 3).
int isw_mesh_trisk(int64_t N, int64_t E, int64_t NV, const double* pos, const int32_t* ring, const uint8_t* deg,
                   const double* corners, const int32_t* eI, const int32_t* eJ, const uint8_t* eKi,
                   const uint8_t* eKj, double* vpos, int32_t* cornerV, int32_t* cellEdge, int8_t* cellSign,
                   int32_t* ev1, int32_t* ev2, double* xe, double* ne, double* xm, double* kite_td, int8_t* circ1,
                   int8_t* circ2, int32_t* cellEdgeAsc, int32_t* vEdge, int8_t* vEdgeSign, int32_t* vCell,
                   int8_t* vSlot) {
  std::unordered_map<uint64_t, int32_t> vid;
  vid.reserve(2 * N);
  int32_t nv = 0;
  for (int64_t i = 0; i < N; i++) {
    const int m = deg[i];
    for (int k = 0; k < 6; k++) {
      if (k >= m) { cornerV[6 * i + k] = -1; continue; }
      int64_t s[3] = {i, ring[6 * i + k], ring[6 * i + (k + 1) % m]};
      std::sort(s, s + 3);
      const uint64_t key = (uint64_t)((s[0] * N + s[1]) * N + s[2]);
      auto it = vid.find(key);
      if (it == vid.end()) {
        vid.emplace(key, nv);
        put3(vpos, nv, get3(corners, 6 * i + k));
        cornerV[6 * i + k] = nv++;
      } else {
        cornerV[6 * i + k] = it->second;
      }
    }
  }
  if (nv != NV) return -1;
  for (int64_t f = 0; f < 6 * N; f++) { cellEdge[f] = -1; cellSign[f] = 0; }
  for (int64_t e = 0; e < E; e++) {
    cellEdge[6 * eI[e] + eKi[e]] = (int32_t)e;
    cellSign[6 * eI[e] + eKi[e]] = 1;
    cellEdge[6 * eJ[e] + eKj[e]] = (int32_t)e;
    cellSign[6 * eJ[e] + eKj[e]] = -1;
  }
#pragma omp parallel for schedule(static)
  for (int64_t e = 0; e < E; e++) {
    const int64_t i = eI[e];
    const int mi = deg[i], ki = eKi[e];
    ev1[e] = cornerV[6 * i + (ki + mi - 1) % mi];
    ev2[e] = cornerV[6 * i + ki];
    const V3 pi = get3(pos, i), pj = get3(pos, eJ[e]);
    const V3 s = add(pi, pj);
    const V3 x = divide(s, norm(s));
    V3 d = sub(pj, pi);
    d = sub(d, scale(x, dot(d, x)));
    const V3 n = divide(d, norm(d));
    put3(xe, e, x);
    put3(ne, e, n);
    const V3 m = add(get3(vpos, ev1[e]), get3(vpos, ev2[e]));
    put3(xm, e, divide(m, norm(m)));
    for (int side = 0; side < 2; side++) {
      const V3 v = get3(vpos, side == 0 ? ev1[e] : ev2[e]);
      const double c = dot(n, cross(x, sub(x, v)));
      (side == 0 ? circ1 : circ2)[e] = (int8_t)(c > 0 ? 1 : c < 0 ? -1 : 0);
    }
  }
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < N; i++) {
    const int m = deg[i];
    const V3 p = get3(pos, i);
    for (int k = 0; k < 6; k++) {
      double* td = kite_td + 4 * (6 * i + k);
      if (k >= m) { td[0] = td[1] = td[2] = td[3] = 0.0; continue; }
      const V3 v = get3(vpos, cornerV[6 * i + k]);
      tri_parts(p, get3(xe, cellEdge[6 * i + k]), v, td[0], td[1]);
      tri_parts(p, v, get3(xe, cellEdge[6 * i + (k + 1) % m]), td[2], td[3]);
    }
    // a cell's edges ascending (pads last)
    int32_t a[6];
    for (int k = 0; k < 6; k++) a[k] = cellEdge[6 * i + k];
    std::stable_sort(a, a + 6, [](int32_t x, int32_t y) {
      return (x < 0 ? INT64_MAX : (int64_t)x) < (y < 0 ? INT64_MAX : (int64_t)y);
    });
    for (int k = 0; k < 6; k++) cellEdgeAsc[6 * i + k] = a[k];
  }
  // a vertex's three edges ascending with their signs; its three (cell, slot) kites by ascending cell
  std::vector<int32_t> vn(NV, 0);
  for (int64_t e = 0; e < E; e++)
    for (int side = 0; side < 2; side++) {
      const int32_t v = side == 0 ? ev1[e] : ev2[e];
      if (vn[v] >= 3) return -2;
      vEdge[3 * v + vn[v]] = (int32_t)e;
      vEdgeSign[3 * v + vn[v]] = side == 0 ? circ1[e] : circ2[e];
      vn[v]++;
    }
  std::fill(vn.begin(), vn.end(), 0);
  for (int64_t i = 0; i < N; i++)
    for (int k = 0; k < deg[i]; k++) {
      const int32_t v = cornerV[6 * i + k];
      if (vn[v] >= 3) return -3;
      vCell[3 * v + vn[v]] = (int32_t)i;
      vSlot[3 * v + vn[v]] = (int8_t)k;
      vn[v]++;
    }
  return 0;
}

// Stage 4: the tangential weights (Thuburn et al. 2009) from the kites (N, 6) and cell areas: for each edge, around
// cell eI then cell eJ, each other edge counter-clockwise from it, weight side (1/2 - R) sign(cell, edge') l_e' / d_e
// with R the running kite fraction of the vertices passed so far; side +1 for eI, -1 for eJ. eoeStart (E + 1), eoe,
// w (sum over edges of deg(eI) - 1 + deg(eJ) - 1).
int isw_mesh_weights(int64_t N, int64_t E, const uint8_t* deg, const int32_t* eI, const int32_t* eJ,
                     const uint8_t* eKi, const uint8_t* eKj, const int32_t* cellEdge, const int8_t* cellSign,
                     const double* kite, const double* areaCell, const double* le, const double* de,
                     int32_t* eoeStart, int32_t* eoe, double* w) {
  eoeStart[0] = 0;
  for (int64_t e = 0; e < E; e++) eoeStart[e + 1] = eoeStart[e] + (deg[eI[e]] - 1) + (deg[eJ[e]] - 1);
#pragma omp parallel for schedule(static)
  for (int64_t e = 0; e < E; e++) {
    int64_t o = eoeStart[e];
    for (int side = 0; side < 2; side++) {
      const int64_t c = side == 0 ? eI[e] : eJ[e];
      const int kc = side == 0 ? eKi[e] : eKj[e];
      const double sg = side == 0 ? 1.0 : -1.0;
      const int m = deg[c];
      double run = 0.0;
      for (int step = 1; step < m; step++) {
        const int k = (kc + step) % m;
        run = run + kite[6 * c + (k + m - 1) % m] / areaCell[c];
        const int32_t e2 = cellEdge[6 * c + k];
        eoe[o] = e2;
        w[o] = sg * (0.5 - run) * cellSign[6 * c + k] * le[e2] / de[e];
        o++;
      }
    }
  }
  return 0;
}

}  // extern "C"
