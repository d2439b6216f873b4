// Lloyd's iteration toward the spherical centroidal Voronoi tessellation on the device (reference:
// src/insilica/grid/icosa.py:194-224, run on the GPU there from level 6, grid/__init__.py:26,95-99): each free
// generator moves to the area-weighted mean of its Voronoi polygon's fan triangles (generator, corner k, corner k + 1),
// every corner the circumcentre of (generator, ring k, ring k + 1) from the previous iterate; generators 0 and 11 (the
// icosahedron's poles) stay; until the largest move (radians) is below tol.
//
// One thread per cell, the iterate double-buffered; each launch is one iteration and folds its largest move into its
// own slot of a chunk of moves (atomicMax on the bits of a non-negative double, which order as the doubles do). A
// launch whose previous slot is below tol returns at once, so the launches queued past convergence do nothing and the
// host finds the converged iteration from the chunk's moves (one read per chunk).
//
// The arithmetic is the host builder's (csrc/mesh.cpp), compiled with -ffp-contract=off: division and sqrt round
// correctly on both sides, and the fan triangle's angle 2 atan2(|t|, d) uses a double-double small-argument
// arctangent correctly rounded for t / d < 1e-2 (a fan triangle's t / d is about 1.6e-3 at level 3 and falls 4x per
// level), the device library's atan2 beyond. The C library's atan2 is not correctly rounded in about one case in a
// million there, so the two builders agree to rounding, as the reference's host and device Lloyd do.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <vector>

#include "insilica_hip.h"

namespace isw_mesh {

constexpr int LLOYD_THREADS = 256;

struct V3 {
  double x, y, z;
};
__device__ inline V3 add(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ inline V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ inline V3 scale(V3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ inline V3 divide(V3 a, double s) { return {a.x / s, a.y / s, a.z / s}; }
__device__ inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ inline V3 cross(V3 a, V3 b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ inline double norm(V3 a) { return sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }
__device__ inline V3 load3(const double* __restrict__ p, int i) { return {p[3 * i], p[3 * i + 1], p[3 * i + 2]}; }

// The tangent pair at unit a: e1 = a x (the coordinate axis least aligned with a, the first of equals), normalised;
// e2 = a x e1.
__device__ inline void frame(V3 a, V3& e1, V3& e2) {
  const double ax = fabs(a.x), ay = fabs(a.y), az = fabs(a.z);
  int k = 0;
  double m = ax;
  if (ay < m) {
    m = ay;
    k = 1;
  }
  if (az < m) k = 2;
  const V3 axis = {k == 0 ? 1.0 : 0.0, k == 1 ? 1.0 : 0.0, k == 2 ? 1.0 : 0.0};
  e1 = cross(a, axis);
  e1 = divide(e1, norm(e1));
  e2 = cross(a, e1);
}

// The unit circumcentre of (a, b, c): the normal of the plane through them from the chords in a's frame.
__device__ inline V3 circumcentre(V3 a, V3 b, V3 c) {
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

// Double-double pieces (Dekker, Knuth): exact sums and products split into a rounded part and its error.
struct DD {
  double hi, lo;
};
__device__ inline DD two_sum(double a, double b) {
  const double s = a + b, bb = s - a;
  return {s, (a - (s - bb)) + (b - bb)};
}
__device__ inline DD quick_sum(double a, double b) {  // |a| >= |b|
  const double s = a + b;
  return {s, b - (s - a)};
}
__device__ inline DD two_prod(double a, double b) {
  const double p = a * b;
  return {p, fma(a, b, -p)};
}
__device__ inline DD dd_mul(DD a, DD b) {
  DD p = two_prod(a.hi, b.hi);
  p.lo += a.hi * b.lo + a.lo * b.hi;
  return quick_sum(p.hi, p.lo);
}
__device__ inline DD dd_add(DD a, DD b) {
  DD s = two_sum(a.hi, b.hi);
  s.lo += a.lo + b.lo;
  return quick_sum(s.hi, s.lo);
}

// atan(y / x) for x > 0, 0 <= y / x < 1e-2, correctly rounded but for cases within ~z^4 ulp of a rounding boundary:
// z = y / x in double-double, atan z = z + z^3 P(z^2) with P the Taylor series -1/3 + z^2/5 - ... to z^10 (the first
// term left out is below 1e-20 z there), its leading -1/3 in double-double.
__device__ inline double atan_small(double y, double x) {
  const double q = y / x;
  const DD z = quick_sum(q, fma(-q, x, y) / x);
  const DD z2 = dd_mul(z, z);
  const double t = z2.hi;
  const double p2 = 1.0 / 5 + t * (-1.0 / 7 + t * (1.0 / 9 + t * (-1.0 / 11 + t * (1.0 / 13))));
  const DD third = {-0.33333333333333331, -1.8503717077085942e-17};
  const DD P = dd_add(third, two_prod(t, p2));
  const DD r = dd_add(z, dd_mul(dd_mul(z, z2), P));
  return r.hi + r.lo;
}

// The spherical triangle (a, b, c)'s area 2 atan2(|t|, d): t = a.(b x c) from the chords in a's frame, d = 1 + a.b +
// b.c + c.a as 4 - (|b - a|^2 + |c - b|^2 + |a - c|^2) / 2.
__device__ inline double tri_area(V3 a, V3 b, V3 c) {
  V3 e1, e2;
  frame(a, e1, e2);
  const V3 u = sub(b, a), v = sub(c, a), w = sub(c, b);
  const double t = fabs(dot(u, e1) * dot(v, e2) - dot(u, e2) * dot(v, e1));
  const double d = 4 - 0.5 * (dot(u, u) + dot(w, w) + dot(v, v));
  return 2 * (d > 0 && t < 1e-2 * d ? atan_small(t, d) : atan2(t, d));
}

// One Jacobi iteration: P the iterate, Q the next; moved[slot] gets the largest move; a launch after a converged
// one (moved[slot - 1] < tol) does nothing.
__global__ __launch_bounds__(LLOYD_THREADS) void k_lloyd(int N, const double* __restrict__ P, double* __restrict__ Q,
                                                         const int32_t* __restrict__ ring,
                                                         const uint8_t* __restrict__ deg,
                                                         unsigned long long* __restrict__ moved, int slot,
                                                         double tol) {
  if (slot > 0 && __longlong_as_double((long long)moved[slot - 1]) < tol) return;
  __shared__ double wmax[LLOYD_THREADS / 64];
  const int i = blockIdx.x * LLOYD_THREADS + threadIdx.x;
  double mv = 0;
  if (i < N) {
    const V3 a = load3(P, i);
    const int m = deg[i] < 6 ? deg[i] : 6;
    V3 C[6];
    for (int k = 0; k < 6; k++)
      if (k < m) C[k] = circumcentre(a, load3(P, ring[6 * i + k]), load3(P, ring[6 * i + (k + 1 < m ? k + 1 : 0)]));
    V3 s = {0, 0, 0};
    for (int k = 0; k < 6; k++)
      if (k < m) {
        const V3 c0 = C[k], c1 = C[k + 1 < m ? k + 1 : 0];
        s = add(s, scale(add(add(a, c0), c1), tri_area(a, c0, c1)));
      }
    V3 p = divide(s, norm(s));
    if (i == 0 || i == 11) {
      p = a;
    } else {
      mv = norm(sub(p, a));
    }
    Q[3 * i] = p.x;
    Q[3 * i + 1] = p.y;
    Q[3 * i + 2] = p.z;
  }
  for (int o = 32; o > 0; o >>= 1) mv = fmax(mv, __shfl_xor(mv, o));
  if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = mv;
  __syncthreads();
  if (threadIdx.x == 0) {
    double b = wmax[0];
    for (int w = 1; w < LLOYD_THREADS / 64; w++) b = fmax(b, wmax[w]);
    atomicMax(&moved[slot], (unsigned long long)__double_as_longlong(b));
  }
}

__global__ void k_zero_u64(unsigned long long* x, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = 0ull;
}

}  // namespace isw_mesh

extern "C" int32_t isw_mesh_lloyd_device(int64_t N, double* pos, const int32_t* ring, const uint8_t* deg, int32_t max_iters,
                                  double tol, double* work, int32_t chunk, int32_t* iters, double* moved,
                                  void* stream) {
  using namespace isw_mesh;
  if (N < 12 || N > (int64_t)1 << 28 || !pos || !ring || !deg || !work || !iters || !moved || max_iters < 1 ||
      chunk < 1 || chunk > ISW_LLOYD_MAX_CHUNK || !(tol > 0))
    return ISW_EBADARG;
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  double* buf[2] = {pos, work};
  unsigned long long* mv = reinterpret_cast<unsigned long long*>(work + 3 * N);
  const int n = (int)N, blocks = (n + LLOYD_THREADS - 1) / LLOYD_THREADS;
  std::vector<unsigned long long> host(chunk);
  int done = 0;
  double last = 0;
  hipError_t e = hipSuccess;
  while (done < max_iters) {
    const int k = max_iters - done < chunk ? max_iters - done : chunk;
    k_zero_u64<<<(k + 255) / 256, 256, 0, st>>>(mv, k);
    for (int s = 0; s < k; s++) {
      const int it = done + s + 1;  // reads buf[(it - 1) & 1], writes buf[it & 1]
      k_lloyd<<<blocks, LLOYD_THREADS, 0, st>>>(n, buf[(it - 1) & 1], buf[it & 1], ring, deg, mv, s, tol);
    }
    if ((e = hipGetLastError()) != hipSuccess) return (int32_t)e;
    if ((e = hipMemcpyAsync(host.data(), mv, sizeof(unsigned long long) * k, hipMemcpyDeviceToHost, st)) !=
        hipSuccess)
      return (int32_t)e;
    if ((e = hipStreamSynchronize(st)) != hipSuccess) return (int32_t)e;
    int conv = -1;
    for (int s = 0; s < k; s++) {
      double v;
      memcpy(&v, &host[s], sizeof v);
      last = v;
      if (v < tol) {
        conv = s;
        break;
      }
    }
    if (conv >= 0) {
      done += conv + 1;
      break;
    }
    done += k;
  }
  // the last iterate to pos
  if (done & 1) {
    if ((e = hipMemcpyAsync(pos, work, sizeof(double) * 3 * N, hipMemcpyDeviceToDevice, st)) != hipSuccess)
      return (int32_t)e;
    if ((e = hipStreamSynchronize(st)) != hipSuccess) return (int32_t)e;
  }
  *iters = done;
  *moved = last;
  return 0;
}
