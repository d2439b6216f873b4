// The TRiSK layered core, FCT transport, SWE solver and air-sea coupling as gfx950 HIP kernels, with the host-side
// sequencing of each C-ABI entry point (the reference's Python drivers: layers.py, transport.py, swe.py, ocean.py).
//
// This file is compiled twice: with ISW_FORM_MIXED=0 and -ffp-contract=off into namespace isw_f64 (the plain FP64 form:
// every operation rounds on its own, as the reference's --fmad=false build), and with ISW_FORM_MIXED=1 and
// -ffp-contract=fast into namespace isw_mix (the reference's mixed form). capi.hip dispatches on the `prec` field.
//
// Design (MI355X): one thread per cell, edge or vertex, 256-thread blocks (four 64-wide waves), layer-major arrays so
// consecutive threads touch consecutive words. The reference's one-kernel-per-loop structure is fused wherever a
// value can be recomputed bit for bit instead of stored and re-read: the Bernoulli (cell) and PV (vertex) passes run
// in one launch; the air fast sub-steps read the stage's start state directly (no copies); the stage flux
// F = h_e * mean(u) is recomputed inside the limiter and the divergence instead of being stored, the limited
// flux's volume Fm is written by each edge's eI cell, and the wall mask is applied inside the kernels that write.
// Every sum keeps the reference's order (a cell's edges in cellEdge or cellEdgeAsc order, a vertex's three in table
// order, an edge's eoe neighbours ascending), so the FP64 form matches the reference's double path value for value.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdlib>

#include "insilica_hip.h"
#include "diag.hpp"

#if ISW_FORM_MIXED
#define NS isw_mix
#define DIVBY(x, r) ((x) * (r))
typedef float fast;
#else
#define NS isw_f64
#define DIVBY(x, r) ((x) / (r))
typedef double fast;
#endif

namespace NS {

constexpr int BS = 256;
#define TID ((int)(blockIdx.x * blockDim.x + threadIdx.x))

static inline dim3 g1(long n) { return dim3((unsigned)((n + BS - 1) / BS)); }
static inline dim3 g2(long n, int y) { return dim3((unsigned)((n + BS - 1) / BS), (unsigned)y); }

// Passes over cells or vertices (40,962 / 81,920 at level 6) launch 64-thread blocks (one wave each) so that they
// still spread over all 256 CUs; column passes launch (64, L) blocks: 64 cells or edges, one thread row per layer.
constexpr int BS1 = 64;
static inline dim3 c1(long n) { return dim3((unsigned)((n + BS1 - 1) / BS1)); }
static inline dim3 c2(long n, int y) { return dim3((unsigned)((n + BS1 - 1) / BS1), (unsigned)y); }
// (64, 4) blocks over cells (or edges) x rows: four rows of the same 64 elements per block, their index tables
// and stencil weights shared in L1
static inline dim3 c4(long n, int y) { return dim3((unsigned)((n + BS1 - 1) / BS1), (unsigned)((y + 3) / 4)); }

// XCD-aware block order: the dispatcher deals consecutive blocks to the 8 XCDs in turn; renumbered so that each XCD
// takes one contiguous range of blocks, neighbouring cells (Hilbert-ordered) then gather through the same L2.
__device__ __forceinline__ int xblock() {
  const int nb = (int)gridDim.x, b = (int)blockIdx.x;
  const int q = nb >> 3, r = nb & 7, x = b & 7, k = b >> 3;
  return (x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + k;
}
#define XTID ((int)(xblock() * blockDim.x + threadIdx.x))

// A cell's six table entries (edges or neighbours; pads -1 last) in three 8-byte loads, so that every gather that
// depends on them can be issued at once; sums then skip the pads in table order, as the reference's loops do.
__device__ __forceinline__ void ring6(const int* __restrict__ tab, int i, int (&e)[6]) {
  const int2* p = reinterpret_cast<const int2*>(tab + (size_t)i * 6);
  const int2 a = p[0], b = p[1], c = p[2];
  e[0] = a.x; e[1] = a.y; e[2] = b.x; e[3] = b.y; e[4] = c.x; e[5] = c.y;
}
__device__ __forceinline__ int ok0(int e) { return e < 0 ? 0 : e; }

// A cell's six per-slot coefficients of one layer (contiguous, 16-byte aligned as doubles, 8-byte as floats), in
// three vector loads.
__device__ __forceinline__ void load6(const double* p, double (&c)[6]) {
  const double2* q = reinterpret_cast<const double2*>(p);
  const double2 a = q[0], b = q[1], d = q[2];
  c[0] = a.x; c[1] = a.y; c[2] = b.x; c[3] = b.y; c[4] = d.x; c[5] = d.y;
}
__device__ __forceinline__ void load6(const float* p, double (&c)[6]) {
  const float2* q = reinterpret_cast<const float2*>(p);
  const float2 a = q[0], b = q[1], d = q[2];
  c[0] = a.x; c[1] = a.y; c[2] = b.x; c[3] = b.y; c[4] = d.x; c[5] = d.y;
}

// Settings of a layer stack passed by value (reference: Layers.__init__, src/insilica/layers.py:161-218).
struct LP {
  int L, air;
  double grav, cp, kappa, p0, floor_, div_damp, thin, shelf, nu4;
  double gp[ISW_LMAX], drag[ISW_LMAX], stress[ISW_LMAX];
  // a batch of worlds (all NULL for one world): the world of each cell and edge, per-world gravity, del^4 coefficient,
  // fast sub-steps, their length and 1/n per stage (isw_batch)
  const int32_t *wc, *we, *wn;
  const double *wgrav, *wnu4, *wtau, *winv;
  // tides (Boussinesq layers; NULL / 0 without): the potential per cell, the internal tide's drag per edge, the
  // bottom's quadratic drag coefficient
  const double *tide, *cbot;
  double cd_bot;
};

static LP make_lp(const isw_layers* s) {
  LP p;
  p.L = s->L; p.air = s->air;
  p.grav = s->grav; p.cp = s->cp; p.kappa = s->kappa; p.p0 = s->p0; p.floor_ = s->floor_; p.div_damp = s->div_damp;
  p.thin = s->thin; p.shelf = s->shelf; p.nu4 = s->nu4;
  for (int k = 0; k < ISW_LMAX; k++) { p.gp[k] = s->gp[k]; p.drag[k] = s->drag[k]; p.stress[k] = s->stress[k]; }
  const isw_batch* b = s->batch;
  p.wc = b ? b->wc : nullptr; p.we = b ? b->we : nullptr; p.wn = b ? b->nsub : nullptr;
  p.wgrav = b ? b->grav : nullptr; p.wnu4 = b ? b->nu4 : nullptr; p.wtau = b ? b->tau : nullptr;
  p.winv = b ? b->inv_n : nullptr;
  p.tide = s->air ? nullptr : s->tide;
  p.cbot = s->air ? nullptr : s->cbot;
  p.cd_bot = s->air ? 0.0 : s->cd_bot;
  return p;
}

// Gravity at a cell and the del^4 coefficient at an edge: the world's in a batch, else the stack's.
__device__ __forceinline__ double grav_at(const LP& p, int i) { return p.wc ? p.wgrav[p.wc[i]] : p.grav; }
__device__ __forceinline__ double nu4_at(const LP& p, int e) { return p.we ? p.wnu4[p.we[e]] : p.nu4; }

// A fast sub-step's element: in a batch its world's sub-step count and length for this stage (a world whose
// sub-steps are done skips), else the stack's.
struct Sub {
  bool active, last;
  double tau;
};
__device__ __forceinline__ Sub sub_at(const LP& p, const int32_t* wid, int x, int stage, int it, int n, double tau) {
  if (wid) {
    const int w = wid[x];
    n = p.wn[w * 3 + stage];
    tau = p.wtau[w * 3 + stage];
  }
  return Sub{it < n, it == n - 1, tau};
}

// What the DIVBY divisions take: the quantity (plain form) or its stored reciprocal (mixed form).
struct Div {
  const double *A, *de, *AV;
};
static Div divs(const isw_grid* g) {
#if ISW_FORM_MIXED
  return Div{g->rA, g->rde, g->rAV};
#else
  return Div{g->areaCell, g->de, g->areaV};
#endif
}

__global__ void k_copy(long n, const double* __restrict__ a, double* __restrict__ b) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = a[i];
}

__global__ void k_fill(long n, double v, double* __restrict__ x) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = v;
}

// A layer's thickness at an edge (layers.cu:42-45), ob NULL.
__device__ __forceinline__ double edge_h(double a, double b, double shelf) {
  const double m = a < b ? a : b;
  return shelf > 0 && m < shelf ? m : 0.5 * (a + b);
}

// A cell's kinetic energy times its area, sum over its edges of l d u^2 / 4, in table order.
__device__ __forceinline__ double cell_ke(const isw_grid& g, const double* __restrict__ ul, const int (&e)[6]) {
  double t[6];
#pragma unroll
  for (int m = 0; m < 6; m++) {
    const int ek = ok0(e[m]);
    t[m] = g.le[ek] * g.de[ek] * 0.25 * ul[ek] * ul[ek];
  }
  double ke = 0;
#pragma unroll
  for (int m = 0; m < 6; m++)
    if (e[m] >= 0) ke += t[m];
  return ke;
}

// ---------------------------------------------------------------------------------------------------------------
// Mesh operators: the vector Laplacian (trisk.cu:14-45). Cells and vertices in one launch.
// ---------------------------------------------------------------------------------------------------------------
__global__ __launch_bounds__(BS) void k_div_curl(isw_grid g, int L, int nbc, const double* __restrict__ u,
                                                 double* __restrict__ D, double* __restrict__ Z) {
  if ((int)blockIdx.x < nbc) {
    const int t = TID;
    if (t >= g.N * L) return;
    const int l = t / g.N, i = t - l * g.N;
    const double* ul = u + (size_t)l * g.E;
    int e[6];
    ring6(g.cellEdge, i, e);
    double tm[6];
#pragma unroll
    for (int k = 0; k < 6; k++) tm[k] = (double)g.cellSign[i * 6 + k] * g.le[ok0(e[k])] * ul[ok0(e[k])];
    double s = 0;
#pragma unroll
    for (int k = 0; k < 6; k++)
      if (e[k] >= 0) s += tm[k];
    D[t] = s / g.areaCell[i];
  } else {
    const int t = (int)((blockIdx.x - nbc) * blockDim.x + threadIdx.x);
    if (t >= g.NV * L) return;
    const int l = t / g.NV, v = t - l * g.NV;
    const double* ul = u + (size_t)l * g.E;
    double z = 0;
#pragma unroll
    for (int k = 0; k < 3; k++) {
      const int e = g.vEdge[v * 3 + k];
      z += (double)g.vEdgeSign[v * 3 + k] * g.de[e] * ul[e];
    }
    Z[t] = z / g.areaV[v];
  }
}

__global__ __launch_bounds__(BS) void k_edge_lap(isw_grid g, int L, const double* __restrict__ D,
                                                 const double* __restrict__ Z, double* __restrict__ out) {
  const int t = XTID;
  if (t >= g.E * L) return;
  const int l = t / g.E, e = t - l * g.E;
  const double* Dl = D + (size_t)l * g.N;
  const double* Zl = Z + (size_t)l * g.NV;
  out[t] = (Dl[g.eJ[e]] - Dl[g.eI[e]]) / g.de[e] - (Zl[g.ev2[e]] - Zl[g.ev1[e]]) / g.le[e];
}


// ---------------------------------------------------------------------------------------------------------------
// Column kernels at the step's start (layers.cu:33-74,165-175,234-250).
// ---------------------------------------------------------------------------------------------------------------
__global__ __launch_bounds__(BS) void k_air_stiff(int N, LP p, const double* __restrict__ h,
                                                  const double* __restrict__ th, double* __restrict__ S) {
  const int i = XTID;
  if (i >= N) return;
  const double gr = grav_at(p, i);
  double pr = 0;
  for (int k = 0; k < p.L; k++) {
    pr += gr * h[(size_t)k * N + i];
    const double dth = th[(size_t)k * N + i] - (k + 1 < p.L ? th[(size_t)(k + 1) * N + i] : 0.0);
    const double dpi = p.kappa * p.cp * pow(pr / p.p0, p.kappa) / pr;
    S[(size_t)k * N + i] = dth > 0 ? dth * gr * dpi : 0.0;
  }
}

__global__ __launch_bounds__(BS) void k_air_gfast(int N, LP p, double rgas, const double* __restrict__ h,
                                                  const double* __restrict__ th, double* __restrict__ Gc) {
  const int i = XTID;
  if (i >= N) return;
  const int L = p.L;
  double mu = 0;
  for (int k = 0; k < L; k++) mu += h[(size_t)k * N + i];
  int kb = L - 1;
  while (kb > 0 && h[(size_t)kb * N + i] < 2 * p.floor_) kb--;
  const double gr = grav_at(p, i);
  const double Ts = th[(size_t)kb * N + i] * pow(gr * mu / p.p0, p.kappa);
  double gc = rgas * Ts / mu;
  double above = 0;
  for (int j = 0; j < L - 1; j++) {
    above += h[(size_t)j * N + i];
    const double dth = th[(size_t)j * N + i] - th[(size_t)(j + 1) * N + i];
    if (dth > 0) gc += dth * rgas * pow(gr * above / p.p0, p.kappa) * above / (mu * mu);
  }
  Gc[i] = gc;
}

__global__ __launch_bounds__(BS) void k_lay_eta(int N, int L, const double* __restrict__ h,
                                                const double* __restrict__ b, double* __restrict__ eta) {
  const int i = XTID;
  if (i >= N) return;
  double z = b[i];
  for (int k = L - 1; k >= 0; k--) z += h[(size_t)k * N + i];
  eta[i] = z;
}

__global__ __launch_bounds__(BS) void k_lay_ubar(isw_grid g, int L, double shelf, const double* __restrict__ h,
                                                 const double* __restrict__ u, const double* __restrict__ wall,
                                                 double* __restrict__ ubar) {
  const int e = XTID;
  if (e >= g.E) return;
  const int N = g.N, eI = g.eI[e], eJ = g.eJ[e];
  double H = 0, U = 0;
  for (int k = 0; k < L; k++) {
    const double he = edge_h(h[(size_t)k * N + eI], h[(size_t)k * N + eJ], shelf);
    H += he;
    U += he * u[(size_t)k * g.E + e];
  }
  ubar[e] = wall && wall[e] == 0 ? 0.0 : U / H;
}

// ---------------------------------------------------------------------------------------------------------------
// Per stage, pass 1: the Bernoulli/Montgomery potential per cell (air_bern / lay_bern, layers.cu:83-158) and the PV
// at vertices (swe_vertex, fused.cu:82-98), in one launch of (64, L) blocks: blocks [0, nbc) take 64 cells, the rest
// 64 vertices; thread row y is layer y. For air each thread takes one interface's Exner power (the expensive `pow`),
// shared with the column through LDS; every thread then replays the column's sums in the reference's order.
// ---------------------------------------------------------------------------------------------------------------
__global__ __launch_bounds__(512) void k_bern_pv(isw_grid g, LP p, Div dv, int nbc, const double* __restrict__ h,
                                                 const double* __restrict__ u, const double* __restrict__ th,
                                                 const double* __restrict__ b, const double* __restrict__ S,
                                                 const double* __restrict__ fV, double* __restrict__ B,
                                                 double* __restrict__ pibar, double* __restrict__ q) {
  const int N = g.N, E = g.E, L = p.L;
  const int j = threadIdx.x, k = threadIdx.y;
  if ((int)blockIdx.x < nbc) {
    __shared__ double sx[ISW_LMAX][BS1];
    const int i = blockIdx.x * BS1 + j;
    const bool live = i < N;
    const int ic = live ? i : 0;
    int e[6];
    ring6(g.cellEdge, ic, e);
    const double ke = cell_ke(g, u + (size_t)k * E, e);
    if (p.air) {
      const double gr = grav_at(p, ic);
      double pr = 0, pt = 0;
      for (int m = 0; m <= k; m++) {
        pt = pr;
        pr = pr + gr * h[(size_t)m * N + ic];
      }
      const double x = pow(pr / p.p0, p.kappa);
      sx[k][j] = x;
      __syncthreads();
      if (!live) return;
      const double xt = k ? sx[k - 1][j] : 0.0;
      pibar[(size_t)k * N + i] = p.cp * (pr * x - pt * xt) / ((p.kappa + 1) * (pr - pt));
      double M = gr * b[i] + th[(size_t)(L - 1) * N + i] * p.cp * sx[L - 1][j];
      for (int m = L - 2; m >= k; m--) M = M + (th[(size_t)m * N + i] - th[(size_t)(m + 1) * N + i]) * p.cp * sx[m][j];
      double P[ISW_LMAX];
      double a = 0;
      for (int m = 0; m < L; m++) { a += h[(size_t)m * N + i]; P[m] = a; }
      double ph = 0;
      for (int m = L - 1; m >= k; m--) ph += S[(size_t)m * N + i] * P[m];
      B[(size_t)k * N + i] = M - ph + DIVBY(ke, dv.A[i]);
    } else {
      if (!live) return;
      double z[ISW_LMAX];
      double s = b[i];
      for (int m = L - 1; m >= 0; m--) { s += h[(size_t)m * N + i]; z[m] = s; }
      double Mp = 0;
      for (int m = 1; m <= k; m++) Mp += p.gp[m - 1] * z[m];
      double Bk = Mp + ke / g.areaCell[i];
      if (p.tide) Bk -= p.tide[i];                 // (the tide's force grad V on every layer, layers.cu:106)
      B[(size_t)k * N + i] = Bk;
    }
  } else {
    const int v = ((int)blockIdx.x - nbc) * BS1 + j;
    if (v >= g.NV) return;
    const int l = k;
    const double* hl = h + (size_t)l * N;
    const double* ul = u + (size_t)l * E;
    double hv = 0, z = 0;
#pragma unroll
    for (int m = 0; m < 3; m++) {
      const int c = g.vCell[v * 3 + m];
      hv += g.kite[c * 6 + g.vSlot[v * 3 + m]] * hl[c] * 1.0;
    }
#pragma unroll
    for (int m = 0; m < 3; m++) {
      const int ee = g.vEdge[v * 3 + m];
      z += (double)g.vEdgeSign[v * 3 + m] * g.de[ee] * ul[ee];
    }
    q[(size_t)l * g.NV + v] = (DIVBY(z, dv.AV[v]) + fV[v]) / DIVBY(hv, dv.AV[v]);
  }
}

// Per stage, pass 2 (lay_edge_hq, layers.cu:254-263): each edge's h_e and q_e; [edge][layer][h_e, u, q_e] packed for
// the PV flux's neighbour gathers.
__global__ __launch_bounds__(BS) void k_edge_hq(isw_grid g, int L, double shelf, const double* __restrict__ h,
                                                const double* __restrict__ q, const double* __restrict__ u,
                                                double* __restrict__ he, fast* __restrict__ pk,
                                                fast* __restrict__ cf) {
  const int t = XTID;
  if (t >= g.E * L) return;
  const int l = t / g.E, e = t - l * g.E;
  const int N = g.N, NV = g.NV;
  const double hv = edge_h(h[(size_t)l * N + g.eI[e]], h[(size_t)l * N + g.eJ[e]], shelf);
  const double qv = 0.5 * (q[(size_t)l * NV + g.ev1[e]] + q[(size_t)l * NV + g.ev2[e]]);
  he[t] = hv;
  if (cf) {
    // the air's fast sub-steps' coefficient sign * l_e * h_e at this edge's slot in each of its two cells (the
    // product the mass step forms, in its order: (sign * l_e) * h_e)
    const int a = g.eI[e], c = g.eJ[e], sa = g.eslot[2 * e], sc = g.eslot[2 * e + 1];
    const double le = g.le[e];
    cf[((size_t)l * N + a) * 6 + sa] = (fast)((double)g.cellSign[a * 6 + sa] * le * hv);
    cf[((size_t)l * N + c) * 6 + sc] = (fast)((double)g.cellSign[c * 6 + sc] * le * hv);
  }
  fast* pp = pk + ((size_t)e * L + l) * 3;
  pp[0] = (fast)hv; pp[1] = (fast)u[t]; pp[2] = (fast)qv;
}

// Per stage, pass 3 (lay_slow, layers.cu:269-359): the slow momentum tendency of every layer of an edge (W-weight PV
// flux with the thin-layer cap, -grad B, drag, exact-exponential stress, pibar grad theta for air, del^4), its
// thickness-weighted column mean G, T' = T - G, H_e and the column wind. The ~10 eoe neighbours' indices and
// weights are loaded before the sum so that their packed records are gathered together.
// One thread per edge with the layers in registers (L a compile-time count): the edge's neighbour list and weights
// load once for every layer, each neighbour's L packed records are one contiguous run, and the column's blend and
// mean need no LDS; the column sums run in the reference's layer order.
template <int L>
__global__ __launch_bounds__(BS) void k_slow_e(isw_grid g, LP p, Div dv, double dts, const fast* __restrict__ pk,
                                               const fast* __restrict__ wf, const double* __restrict__ hedge,
                                               const double* __restrict__ u, const double* __restrict__ B,
                                               const double* __restrict__ lap4, const double* __restrict__ th,
                                               const double* __restrict__ pibar, double* __restrict__ Tp,
                                               double* __restrict__ G, double* __restrict__ He,
                                               double* __restrict__ ubs) {
  const int e = XTID;
  if (e >= g.E) return;
  const int N = g.N, E = g.E;
  const int eI = g.eI[e], eJ = g.eJ[e];
  const int m0 = g.eoeStart[e], m1 = g.eoeStart[e + 1];
  constexpr int NB = 10;
  int nb[NB];
  fast w[NB];
#pragma unroll
  for (int m = 0; m < NB; m++) {
    const int mm = m0 + m < m1 ? m0 + m : (m1 > m0 ? m0 : 0);
    nb[m] = g.eoe[mm];
    w[m] = wf[mm];
  }
  const double rde = dv.de[e];
  const double nu4 = lap4 ? nu4_at(p, e) : 0.0;
  double dk[L], hk[L], uk[L], sqk[L];
#pragma unroll
  for (int k = 0; k < L; k++) {
    const size_t o = (size_t)k * E + e;
    const double he = hedge[o];
    const double ue = u[o];
    const fast cap = (fast)(p.thin * he), qf = pk[((size_t)e * L + k) * 3 + 2];
    fast Q = 0;
#pragma unroll
    for (int m = 0; m < NB; m++) {
      if (m0 + m >= m1) break;
      const fast* pp = pk + ((size_t)nb[m] * L + k) * 3;
      fast h2 = pp[0];
      if (h2 > cap) h2 = cap;
      const fast F2 = h2 * pp[1];
      Q += w[m] * F2 * (fast)0.5 * (qf + pp[2]);
    }
    for (int m = m0 + NB; m < m1; m++) {           // (meshes with more neighbours per edge than 10)
      const fast* pp = pk + ((size_t)g.eoe[m] * L + k) * 3;
      fast h2 = pp[0];
      if (h2 > cap) h2 = cap;
      const fast F2 = h2 * pp[1];
      Q += wf[m] * F2 * (fast)0.5 * (qf + pp[2]);
    }
    const double shk = p.shelf;
    const double sq0 =
        shk > 0 && he < shk ? (he - 2 * p.floor_ > 0 ? (he - 2 * p.floor_) / (shk - 2 * p.floor_) : 0.0) : 1.0;
    const double sqv = th ? 1.0 : sq0;
    const double* Bl = B + (size_t)k * N;
    double d = Q - DIVBY(sqv * (Bl[eJ] - Bl[eI]), rde) - (p.drag[k] + (1 - sqv) / 21600.0) * ue;
    double ck = p.stress[k];
    if (k == L - 1 && (p.cd_bot > 0 || p.cbot)) {
      // (tides, layers.cu:305-315) the bottom layer's quadratic stress cd |u|, |u| from the normal wind and the TRiSK
      // tangential wind of the packed neighbours' u; the internal tide's drag
      if (p.cd_bot > 0) {
        fast vt = 0;
#pragma unroll
        for (int m = 0; m < NB; m++) {
          if (m0 + m >= m1) break;
          vt += w[m] * pk[((size_t)nb[m] * L + k) * 3 + 1];
        }
        for (int m = m0 + NB; m < m1; m++) vt += wf[m] * pk[((size_t)g.eoe[m] * L + k) * 3 + 1];
        ck += p.cd_bot * sqrt(ue * ue + (double)vt * (double)vt);
      }
      if (p.cbot) ck += p.cbot[e];
    }
    if (ck != 0) {
      const double r = ck / (he > p.floor_ ? he : p.floor_);
      d = d - ((r * dts > 1e-8) ? (1 - exp(-r * dts)) / dts : r) * ue;
    }
    if (th) {
      const double* tl = th + (size_t)k * N;
      const double* pl = pibar + (size_t)k * N;
      d = d + DIVBY(0.5 * (pl[eI] + pl[eJ]) * (tl[eJ] - tl[eI]), rde);
    }
    if (lap4) d = d - nu4 * lap4[o];
    dk[k] = d;
    hk[k] = he;
    uk[k] = ue;
    sqk[k] = sq0;
  }
  // the blend of a thin layer's tendency with its nearest thick neighbours' (shelf), then the column mean
  const bool blend = th && p.shelf > 0;
  auto blended = [&](int m) {
    double Tm = dk[m];
    if (blend && sqk[m] < 1) {
      double wgt = 0, s = 0;
      for (int q2 = m - 1; q2 >= 0; q2--) if (sqk[q2] >= 1) { wgt += hk[q2]; s += hk[q2] * dk[q2]; break; }
      for (int q2 = m + 1; q2 < L; q2++) if (sqk[q2] >= 1) { wgt += hk[q2]; s += hk[q2] * dk[q2]; break; }
      if (wgt > 0) Tm = sqk[m] * dk[m] + (1 - sqk[m]) * s / wgt;
    }
    return Tm;
  };
  double H = 0, U = 0, GT = 0;
#pragma unroll
  for (int m = 0; m < L; m++) {
    H += hk[m];
    U += hk[m] * uk[m];
  }
  double Tb[L];
#pragma unroll
  for (int m = 0; m < L; m++) {
    Tb[m] = blended(m);
    GT += hk[m] * Tb[m];
  }
  const double gg = GT / H;
#pragma unroll
  for (int k = 0; k < L; k++) Tp[(size_t)k * E + e] = Tb[k] - gg;
  G[e] = gg;
  He[e] = H;
  ubs[e] = U / H;
}

// ---------------------------------------------------------------------------------------------------------------
// Air fast sub-steps (layers.cu:182-224): mass then wind, forward-backward, from the stage's start.
// The first sub-step reads the step's start (mu from h0, u from u0) and starts acc; the last writes the stage's wind.
// ---------------------------------------------------------------------------------------------------------------
// The mass sub-step: (64, L) blocks, thread (x, y) updates layer y's mass of one of 64 cells (its six edges gathered
// at once); the column's forward-weighted masses and the linear potential Phi* are then summed through LDS in the
// reference's order (P from the top, Phi from the bottom) by every thread for its own layer.
__global__ __launch_bounds__(512) void k_af_muphi(isw_grid g, LP p, Div dv, int stage, int it, int n, double tau,
                                                  const double* __restrict__ hedge,
                                                  const double* __restrict__ uf, const double* mu_src, double* mu,
                                                  const double* __restrict__ S, double* __restrict__ phi) {
  __shared__ double sa[ISW_LMAX][BS1], sS[ISW_LMAX][BS1];
  const int N = g.N, E = g.E, L = p.L;
  const int j = threadIdx.x, k = threadIdx.y;
  const int i = xblock() * BS1 + j;
  const bool live = i < N;
  const int ic = live ? i : 0;
  const Sub sb = sub_at(p, p.wc, ic, stage, it, n, tau);
  int e[6];
  ring6(g.cellEdge, ic, e);
  const double* hl = hedge + (size_t)k * E;
  const double* ul = uf + (size_t)k * E;
  double t[6];
#pragma unroll
  for (int m = 0; m < 6; m++) {
    const int ek = ok0(e[m]);
    t[m] = (double)g.cellSign[ic * 6 + m] * g.le[ek] * hl[ek] * ul[ek];
  }
  double div = 0;
#pragma unroll
  for (int m = 0; m < 6; m++)
    if (e[m] >= 0) div += t[m];
  const size_t o = (size_t)k * N + ic;
  const double m0 = mu_src[o];
  const double m1 = m0 - DIVBY(sb.tau * div, dv.A[ic]);
  sa[k][j] = m1 + p.div_damp * (m1 - m0);
  sS[k][j] = S[o];
  __syncthreads();
  if (!live || !sb.active) return;
  mu[o] = m1;
  double P[ISW_LMAX];
  double a = 0;
  for (int m = 0; m < L; m++) { a += sa[m][j]; P[m] = a; }
  double ph = 0;
  for (int m = L - 1; m >= k; m--) ph += sS[m][j] * P[m];
  phi[o] = ph;
}

__global__ __launch_bounds__(BS) void k_af_u(isw_grid g, int L, Div dv, LP p, int stage, int it, int n, double tau,
                                             const double* __restrict__ phi, const double* __restrict__ Tp,
                                             const double* __restrict__ G, const double* uin, double* acc, double* uf,
                                             double* un) {
  const int t = XTID;
  if (t >= g.E * L) return;
  const int l = t / g.E, e = t - l * g.E;
  const Sub sb = sub_at(p, p.we, e, stage, it, n, tau);
  if (!sb.active) return;
  const double v = uin[t];
  acc[t] = it == 0 ? 0.0 + v : acc[t] + v;
  const double* pl = phi + (size_t)l * g.N;
  (sb.last ? un : uf)[t] = v + sb.tau * (DIVBY(-(pl[g.eJ[e]] - pl[g.eI[e]]), dv.de[e]) + Tp[t] + G[e]);
}

// One air fast sub-step in one launch, on patches of consecutive (Hilbert-ordered, so compact) cells: each block
// takes the mass step and Phi* of its patch's cells and of the halo cells across its edges (the halo's recomputed
// from the same inputs, so every copy is the same to the bit), keeps Phi* in LDS, then steps the wind on the edges
// whose first cell is in the patch. The arithmetic is k_af_muphi's and k_af_u's in their order. Masses and winds
// alternate between two buffers (a neighbour block may still read this sub-step's inputs), so in/out never alias.
// Patch tables (isw_grid): patch b holds cells [b pcells, (b + 1) pcells), its halo cells ph_cell[ph_off[b] ..],
// its edges pe_edge[pe_off[b] ..] with their cells' LDS slots pe_slot (first | second << 16; halo slots from pcells).
constexpr int FBS = ISW_PAIR_THREADS;
template <int L>
__global__ __launch_bounds__(FBS) void k_af_patch(isw_grid g, LP p, Div dv, int stage, int it, int n, double tau,
                                                  const fast* __restrict__ coef, const double* __restrict__ uin,
                                                  const double* __restrict__ mu_src, double* __restrict__ mu_out,
                                                  const double* __restrict__ S, const double* __restrict__ Tp,
                                                  const double* __restrict__ G, double* __restrict__ acc,
                                                  double* __restrict__ uout, double* __restrict__ un) {
  extern __shared__ double sphi[];                 // [L][pslots]
  const int N = g.N, E = g.E, PC = g.pcells, PS = g.pslots;
  const int b = xblock();
  const int c0 = b * PC, nown = min(PC, N - c0);
  const int h0 = g.ph_off[b], nh = g.ph_off[b + 1] - h0;
  const int e0 = g.pe_off[b], ne = g.pe_off[b + 1] - e0;
  // the wind step's inputs for this thread's first edge, loaded before the mass step so their latency overlaps it
  // (they do not depend on it); further edges, if a patch has more than FBS, load in the loop below
  const int q0 = threadIdx.x;
  int e_ = 0, sl_ = 0;
  Sub sbe{false, false, 0.0};
  double v_[L], a_[L], tp_[L], rde_ = 0, G_ = 0;
  if (q0 < ne) {
    e_ = g.pe_edge[e0 + q0];
    sl_ = g.pe_slot[e0 + q0];
    sbe = sub_at(p, p.we, e_, stage, it, n, tau);
    if (sbe.active) {
      rde_ = dv.de[e_];
      G_ = G[e_];
#pragma unroll
      for (int l = 0; l < L; l++) {
        const size_t t = (size_t)l * E + e_;
        v_[l] = uin[t];
        a_[l] = it == 0 ? 0.0 : acc[t];
        tp_[l] = Tp[t];
      }
    }
  }
  for (int s = threadIdx.x; s < nown + nh; s += FBS) {
    const int i = s < nown ? c0 + s : g.ph_cell[h0 + s - nown];
    const int slot = s < nown ? s : PC + (s - nown);
    const Sub sb = sub_at(p, p.wc, i, stage, it, n, tau);
    if (!sb.active) continue;
    int e[6];
    ring6(g.cellEdge, i, e);
    double m1k[L], sak[L];
#pragma unroll
    for (int k = 0; k < L; k++) {
      double c6[6];                                // (sign * l_e * h_e per slot, written with the stage's h_e)
      load6(coef + ((size_t)k * N + i) * 6, c6);
      const double* ul = uin + (size_t)k * E;
      double t[6];
#pragma unroll
      for (int m = 0; m < 6; m++) t[m] = c6[m] * ul[ok0(e[m])];
      double div = 0;
#pragma unroll
      for (int m = 0; m < 6; m++)
        if (e[m] >= 0) div += t[m];
      const double m0 = mu_src[(size_t)k * N + i];
      const double m1 = m0 - DIVBY(sb.tau * div, dv.A[i]);
      m1k[k] = m1;
      sak[k] = m1 + p.div_damp * (m1 - m0);
    }
    double P[L], Sk[L];
    double a = 0;
#pragma unroll
    for (int m = 0; m < L; m++) {
      a += sak[m];
      P[m] = a;
      Sk[m] = S[(size_t)m * N + i];
    }
#pragma unroll
    for (int k = 0; k < L; k++) {
      double ph = 0;
#pragma unroll
      for (int m = L - 1; m >= k; m--) ph += Sk[m] * P[m];
      sphi[k * PS + slot] = ph;
      if (s < nown) mu_out[(size_t)k * N + i] = m1k[k];
    }
  }
  __syncthreads();
  if (q0 < ne && sbe.active) {
    const int sI = sl_ & 0xffff, sJ = sl_ >> 16;
    double* out = sbe.last ? un : uout;
#pragma unroll
    for (int l = 0; l < L; l++) {
      const size_t t = (size_t)l * E + e_;
      acc[t] = it == 0 ? 0.0 + v_[l] : a_[l] + v_[l];
      out[t] = v_[l] + sbe.tau * (DIVBY(-(sphi[l * PS + sJ] - sphi[l * PS + sI]), rde_) + tp_[l] + G_);
    }
  }
  for (int q = q0 + FBS; q < ne; q += FBS) {
    const int e = g.pe_edge[e0 + q];
    const Sub sb = sub_at(p, p.we, e, stage, it, n, tau);
    if (!sb.active) continue;
    const int sl = g.pe_slot[e0 + q];
    const int sI = sl & 0xffff, sJ = sl >> 16;
    const double rde = dv.de[e], Ge = G[e];
    double* out = sb.last ? un : uout;
#pragma unroll
    for (int l = 0; l < L; l++) {
      const size_t t = (size_t)l * E + e;
      const double v = uin[t];
      acc[t] = it == 0 ? 0.0 + v : acc[t] + v;
      out[t] = v + sb.tau * (DIVBY(-(sphi[l * PS + sJ] - sphi[l * PS + sI]), rde) + Tp[t] + Ge);
    }
  }
}

// ---------------------------------------------------------------------------------------------------------------
// Ocean barotropic sub-steps (layers.cu:366-388) with the wall mask of the reference's lay_mask calls fused in.
// ---------------------------------------------------------------------------------------------------------------
// Two air fast sub-steps (it, it + 1) in one launch on the grid's patches (tables ph2_*, px_*, pc_*): A) the mass
// step and Phi* of every cell of the patch's edges' cells' edges (patch, halo, the halo's halo); B) the wind of
// sub-step it on every edge of patch and halo, in LDS; C) the mass step and Phi* of sub-step it + 1 on patch and
// halo from B's winds; D) the wind of sub-step it + 1 on the patch's edges. Every value is computed as k_af_muphi and
// k_af_u compute it, so the result is the same to the bit as two sub-steps of one launch each.
// Latency: the kernel is a chain of dependent gathers (table -> edge -> value) between three barriers, with ~2.5
// waves per SIMD at level 6, so only A gathers: thread q loads its own edge's inputs (B, D) and the q-th halo edge's
// (B) before A, and A leaves each patch-and-halo cell's coefficients sign * l_e * h_e per layer in LDS (cf, L <= 4)
// with its stiffness and area in registers (patch and halo come first in A's order, so thread q does cell q in both),
// so that C reads no global memory. Halo edges past the block's FBS threads (a few at level 6) load as they go.
template <int L>
__global__ __launch_bounds__(FBS) void k_af_patch2(isw_grid g, LP p, Div dv, int stage, int it, int n, double tau,
                                                   const fast* __restrict__ coef, const double* __restrict__ uin,
                                                   const double* __restrict__ mu_src, double* __restrict__ mu_out,
                                                   const double* __restrict__ S, const double* __restrict__ Tp,
                                                   const double* __restrict__ G, double* __restrict__ acc,
                                                   double* __restrict__ uout, double* __restrict__ un) {
  constexpr bool CF = L <= 4;                      // C's coefficients from LDS (else C gathers them again)
  extern __shared__ double sm[];
  const int N = g.N, E = g.E, PC = g.pcells, PS = g.pslots2, PE = g.pedges2, P1 = g.pslots;
  double* phiA = sm;                               // [L][PS] Phi* of sub-step it
  double* muA = phiA + L * PS;                     // [L][PS] the masses after it (patch and halo)
  double* uB = muA + L * PS;                       // [L][PE] the winds after it
  double* phiC = uB + L * PE;                      // [L][PS] Phi* of sub-step it + 1
  double* cfs = phiC + L * PS;                     // [L][6][P1] (CF) the coefficients of patch and halo cells
  const int b = xblock();
  const int c0 = b * PC, nown = min(PC, N - c0);
  const int h0 = g.ph_off[b], nh = g.ph_off[b + 1] - h0;
  const int k0 = g.ph2_off[b], nk = g.ph2_off[b + 1] - k0;
  const int e0 = g.pe_off[b], ne = g.pe_off[b + 1] - e0;
  const int x0 = g.px_off[b], nx = g.px_off[b + 1] - x0;
  const int nC = nown + nh;                        // <= P1 <= FBS (checked with the tables on the host)
  const int q0 = threadIdx.x;
  // this thread's own edge (if any): its inputs for B and D
  int eo = 0, slo = 0;
  Sub s1{false, false, 0.0}, s2{false, false, 0.0};
  double v_[L], a_[L], tp_[L], rde_ = 0, G_ = 0;
  if (q0 < ne) {
    eo = g.pe_edge[e0 + q0];
    slo = g.pe_slot[e0 + q0];
    s1 = sub_at(p, p.we, eo, stage, it, n, tau);
    s2 = sub_at(p, p.we, eo, stage, it + 1, n, tau);
    if (s1.active) {
      rde_ = dv.de[eo];
      G_ = G[eo];
#pragma unroll
      for (int l = 0; l < L; l++) {
        const size_t t = (size_t)l * E + eo;
        v_[l] = uin[t];
        a_[l] = it == 0 ? 0.0 : acc[t];
        tp_[l] = Tp[t];
      }
    }
  }
  // the q-th halo edge of B: its inputs
  int xsl = 0;
  Sub sx{false, false, 0.0};
  double xv_[L], xtp_[L], xrde = 0, xG = 0;
  if (q0 < nx) {
    const int xe = g.px_edge[x0 + q0];
    xsl = g.px_slot[x0 + q0];
    sx = sub_at(p, p.we, xe, stage, it, n, tau);
    if (sx.active) {
      xrde = dv.de[xe];
      xG = G[xe];
#pragma unroll
      for (int l = 0; l < L; l++) {
        const size_t t = (size_t)l * E + xe;
        xv_[l] = uin[t];
        xtp_[l] = Tp[t];
      }
    }
  }
  // the q-th patch-and-halo cell's edge slots in B's winds (0xffff: no edge)
  int ces[6] = {0xffff, 0xffff, 0xffff, 0xffff, 0xffff, 0xffff};
  if (q0 < nC) {
    const int* ps_ = g.pc_eslot + 3 * g.pc_off[b] + 3 * q0;
    // (unsigned: a missing edge's 0xffff in the high half makes the word negative as an int32)
    const unsigned pa = (unsigned)ps_[0], pb2 = (unsigned)ps_[1], pcc = (unsigned)ps_[2];
    ces[0] = (int)(pa & 0xffffu); ces[1] = (int)(pa >> 16); ces[2] = (int)(pb2 & 0xffffu);
    ces[3] = (int)(pb2 >> 16); ces[4] = (int)(pcc & 0xffffu); ces[5] = (int)(pcc >> 16);
  }
  // A: masses and Phi* of sub-step it on patch, halo and second halo
  int ci = 0, cslot = 0;                           // (thread q's cell q, for C)
  Sub cc{false, false, 0.0};
  double cS[L], cA = 0;
  for (int s = q0; s < nown + nh + nk; s += FBS) {
    const int i = s < nown ? c0 + s : s < nC ? g.ph_cell[h0 + s - nown] : g.ph2_cell[k0 + s - nC];
    const int slot = s < nown ? s : PC + (s - nown);
    const Sub sb = sub_at(p, p.wc, i, stage, it, n, tau);
    if (s < nC) {
      ci = i;
      cslot = slot;
      cc = sub_at(p, p.wc, i, stage, it + 1, n, tau);
    }
    if (!sb.active) continue;
    int e[6];
    ring6(g.cellEdge, i, e);
    double sak[L], Sk[L];
#pragma unroll
    for (int k = 0; k < L; k++) {
      double c6[6];                                // (sign * l_e * h_e per slot, written with the stage's h_e)
      load6(coef + ((size_t)k * N + i) * 6, c6);
      const double* ul = uin + (size_t)k * E;
      double t[6];
#pragma unroll
      for (int m = 0; m < 6; m++) {
        if (CF && s < nC) cfs[(k * 6 + m) * P1 + s] = c6[m];
        t[m] = c6[m] * ul[ok0(e[m])];
      }
      double div = 0;
#pragma unroll
      for (int m = 0; m < 6; m++)
        if (e[m] >= 0) div += t[m];
      const double m0 = mu_src[(size_t)k * N + i];
      const double m1 = m0 - DIVBY(sb.tau * div, dv.A[i]);
      muA[k * PS + slot] = m1;
      sak[k] = m1 + p.div_damp * (m1 - m0);
      Sk[k] = S[(size_t)k * N + i];
    }
    if (s < nC) {
#pragma unroll
      for (int k = 0; k < L; k++) cS[k] = Sk[k];
      cA = dv.A[i];
    }
    double P[L];
    double a = 0;
#pragma unroll
    for (int m = 0; m < L; m++) {
      a += sak[m];
      P[m] = a;
    }
#pragma unroll
    for (int k = 0; k < L; k++) {
      double ph = 0;
#pragma unroll
      for (int m = L - 1; m >= k; m--) ph += Sk[m] * P[m];
      phiA[k * PS + slot] = ph;
    }
  }
  __syncthreads();
  // B: the winds of sub-step it on the patch's edges (registers) and the halo's
  double u1_[L], acc1_[L];
  if (q0 < ne && s1.active) {
    const int sI = slo & 0xffff, sJ = slo >> 16;
#pragma unroll
    for (int l = 0; l < L; l++) {
      acc1_[l] = it == 0 ? 0.0 + v_[l] : a_[l] + v_[l];
      u1_[l] = v_[l] + s1.tau * (DIVBY(-(phiA[l * PS + sJ] - phiA[l * PS + sI]), rde_) + tp_[l] + G_);
      uB[l * PE + q0] = u1_[l];
    }
    if (s1.last) {                                 // (a world whose last sub-step this is: done after B)
#pragma unroll
      for (int l = 0; l < L; l++) {
        const size_t t = (size_t)l * E + eo;
        un[t] = u1_[l];
        acc[t] = acc1_[l];
      }
    }
  }
  if (q0 < nx && sx.active) {
    const int sI = xsl & 0xffff, sJ = xsl >> 16;
#pragma unroll
    for (int l = 0; l < L; l++)
      uB[l * PE + ne + q0] = xv_[l] + sx.tau * (DIVBY(-(phiA[l * PS + sJ] - phiA[l * PS + sI]), xrde) + xtp_[l] + xG);
  }
  for (int w = q0 + FBS; w < nx; w += FBS) {       // (halo edges past the block's threads)
    const int e = g.px_edge[x0 + w];
    const Sub sb = sub_at(p, p.we, e, stage, it, n, tau);
    if (!sb.active) continue;
    const int sl = g.px_slot[x0 + w];
    const int sI = sl & 0xffff, sJ = sl >> 16;
    const double rde = dv.de[e], Ge = G[e];
#pragma unroll
    for (int l = 0; l < L; l++) {
      const size_t t = (size_t)l * E + e;
      uB[l * PE + ne + w] =
          uin[t] + sb.tau * (DIVBY(-(phiA[l * PS + sJ] - phiA[l * PS + sI]), rde) + Tp[t] + Ge);
    }
  }
  __syncthreads();
  // C: masses and Phi* of sub-step it + 1 on patch and halo (thread q's cell q), from B's winds
  if (q0 < nC && cc.active) {
    double cf[L][6];
    if constexpr (CF) {
#pragma unroll
      for (int k = 0; k < L; k++)
#pragma unroll
        for (int m = 0; m < 6; m++) cf[k][m] = cfs[(k * 6 + m) * P1 + q0];
    } else {
#pragma unroll
      for (int k = 0; k < L; k++) load6(coef + ((size_t)k * N + ci) * 6, cf[k]);
    }
    double m1k[L], sak[L];
#pragma unroll
    for (int k = 0; k < L; k++) {
      double t[6];
#pragma unroll
      for (int m = 0; m < 6; m++) t[m] = cf[k][m] * uB[k * PE + (ces[m] != 0xffff ? ces[m] : 0)];
      double div = 0;
#pragma unroll
      for (int m = 0; m < 6; m++)
        if (ces[m] != 0xffff) div += t[m];
      const double m0 = muA[k * PS + cslot];
      const double m1 = m0 - DIVBY(cc.tau * div, cA);
      m1k[k] = m1;
      sak[k] = m1 + p.div_damp * (m1 - m0);
    }
    double P[L];
    double a = 0;
#pragma unroll
    for (int m = 0; m < L; m++) {
      a += sak[m];
      P[m] = a;
    }
#pragma unroll
    for (int k = 0; k < L; k++) {
      double ph = 0;
#pragma unroll
      for (int m = L - 1; m >= k; m--) ph += cS[m] * P[m];
      phiC[k * PS + cslot] = ph;
      if (q0 < nown) mu_out[(size_t)k * N + ci] = m1k[k];
    }
  }
  __syncthreads();
  // D: the winds of sub-step it + 1 on the patch's edges
  if (q0 < ne && s1.active && s2.active) {
    const int sI = slo & 0xffff, sJ = slo >> 16;
    double* out = s2.last ? un : uout;
#pragma unroll
    for (int l = 0; l < L; l++) {
      const size_t t = (size_t)l * E + eo;
      acc[t] = acc1_[l] + u1_[l];
      out[t] = u1_[l] + s2.tau * (DIVBY(-(phiC[l * PS + sJ] - phiC[l * PS + sI]), rde_) + tp_[l] + G_);
    }
  }
}

// One barotropic sub-step in one launch on the grid's patches (as k_af_patch for the air): the surface height of
// the patch's and halo cells, its potential Gc (eta + gamma (eta - eta_prev)) in LDS, then the barotropic wind on
// the patch's edges. Heights alternate between eta and eta_prev, winds between ubb and ubc; an edge's last sub-step
// writes ub (what the stage's update reads). Arithmetic as k_bt_eta's and k_bt_u's.
__global__ __launch_bounds__(FBS) void k_bt_patch(isw_grid g, Div dv, LP p, int stage, int it, int n, double tau,
                                                  double gamma, const double* __restrict__ He,
                                                  const double* __restrict__ ub_src,
                                                  const double* __restrict__ eta_src, double* __restrict__ eta_out,
                                                  const double* __restrict__ G, const double* __restrict__ Gc,
                                                  const double* __restrict__ wall, double* __restrict__ acc,
                                                  double* __restrict__ ub_out, double* __restrict__ ub_fin) {
  extern __shared__ double sx[];                   // [pslots]
  const int N = g.N, PC = g.pcells;
  const int b = xblock();
  const int c0 = b * PC, nown = min(PC, N - c0);
  const int h0 = g.ph_off[b], nh = g.ph_off[b + 1] - h0;
  const int e0 = g.pe_off[b], ne = g.pe_off[b + 1] - e0;
  const int q0 = threadIdx.x;
  int e_ = 0, sl_ = 0;
  Sub sbe{false, false, 0.0};
  double v_ = 0, a_ = 0, rde_ = 0, G_ = 0, w_ = 1;
  if (q0 < ne) {
    e_ = g.pe_edge[e0 + q0];
    sl_ = g.pe_slot[e0 + q0];
    sbe = sub_at(p, p.we, e_, stage, it, n, tau);
    if (sbe.active) {
      v_ = ub_src[e_];
      a_ = it == 0 ? 0.0 : acc[e_];
      rde_ = dv.de[e_];
      G_ = G[e_];
      if (wall) w_ = wall[e_];
    }
  }
  for (int s = threadIdx.x; s < nown + nh; s += FBS) {
    const int i = s < nown ? c0 + s : g.ph_cell[h0 + s - nown];
    const int slot = s < nown ? s : PC + (s - nown);
    const Sub sb = sub_at(p, p.wc, i, stage, it, n, tau);
    if (!sb.active) continue;
    const double et0 = eta_src[i];
    int e[6];
    ring6(g.cellEdge, i, e);
    double t[6];
#pragma unroll
    for (int m = 0; m < 6; m++) {
      const int ek = ok0(e[m]);
      t[m] = (double)g.cellSign[i * 6 + m] * g.le[ek] * He[ek] * ub_src[ek];
    }
    double div = 0;
#pragma unroll
    for (int m = 0; m < 6; m++)
      if (e[m] >= 0) div += t[m];
    const double et = et0 - DIVBY(sb.tau * div, dv.A[i]);
    sx[slot] = Gc[i] * (et + gamma * (et - et0));
    if (s < nown) eta_out[i] = et;
  }
  __syncthreads();
  for (int q = q0; q < ne; q += FBS) {
    int e = e_, sl = sl_;
    Sub sb = sbe;
    double v = v_, a = a_, rde = rde_, Ge = G_, wl = w_;
    if (q != q0) {                                 // (a patch with more edges than threads)
      e = g.pe_edge[e0 + q];
      sl = g.pe_slot[e0 + q];
      sb = sub_at(p, p.we, e, stage, it, n, tau);
      if (!sb.active) continue;
      v = ub_src[e];
      a = it == 0 ? 0.0 : acc[e];
      rde = dv.de[e];
      Ge = G[e];
      wl = wall ? wall[e] : 1.0;
    }
    if (!sb.active) continue;
    double an = it == 0 ? 0.0 + v : a + v;
    double un = v + sb.tau * (DIVBY(-(sx[sl >> 16] - sx[sl & 0xffff]), rde) + Ge);
    if (wall && wl == 0) { un = 0; an = 0; }
    (sb.last ? ub_fin : ub_out)[e] = un;
    acc[e] = an;
  }
}

__global__ __launch_bounds__(BS) void k_bt_eta(isw_grid g, Div dv, LP p, int stage, int it, int n, double tau,
                                               const double* __restrict__ He, const double* __restrict__ ub,
                                               const double* eta_src, double* eta, double* __restrict__ eta_prev) {
  const int i = XTID;
  if (i >= g.N) return;
  const Sub sb = sub_at(p, p.wc, i, stage, it, n, tau);
  if (!sb.active) return;
  const double e0 = eta_src[i];
  eta_prev[i] = e0;
  int e[6];
  ring6(g.cellEdge, i, e);
  double t[6];
#pragma unroll
  for (int m = 0; m < 6; m++) {
    const int ek = ok0(e[m]);
    t[m] = (double)g.cellSign[i * 6 + m] * g.le[ek] * He[ek] * ub[ek];
  }
  double div = 0;
#pragma unroll
  for (int m = 0; m < 6; m++)
    if (e[m] >= 0) div += t[m];
  eta[i] = e0 - DIVBY(sb.tau * div, dv.A[i]);
}

__global__ __launch_bounds__(BS) void k_bt_u(isw_grid g, Div dv, LP p, int stage, int it, int n, double tau,
                                             double gamma, const double* __restrict__ eta,
                                             const double* __restrict__ eta_prev, const double* __restrict__ G,
                                             const double* __restrict__ Gc, const double* __restrict__ wall,
                                             const double* ub_src, double* ub, double* acc) {
  const int e = XTID;
  if (e >= g.E) return;
  const Sub sb = sub_at(p, p.we, e, stage, it, n, tau);
  if (!sb.active) return;
  const double v = ub_src[e];
  double a = it == 0 ? 0.0 + v : acc[e] + v;
  const int i = g.eI[e], j = g.eJ[e];
  const double ej = Gc[j] * (eta[j] + gamma * (eta[j] - eta_prev[j]));
  const double ei = Gc[i] * (eta[i] + gamma * (eta[i] - eta_prev[i]));
  double un = v + sb.tau * (DIVBY(-(ej - ei), dv.de[e]) + G[e]);
  if (wall && wall[e] == 0) { un = 0; a = 0; }
  ub[e] = un;
  acc[e] = a;
}

// ---------------------------------------------------------------------------------------------------------------
// The stage's layer fluxes, limiter and update (layers.cu:220-224,392-447). F is recomputed where it is read:
// air F = h_e acc / n (af_flux); ocean F = h_e (u - ubar*) + h_e ubar_avg (lay_flux), 0 on walls.
// ---------------------------------------------------------------------------------------------------------------
struct FluxSrc {
  int air, stage;
  double inv_n, shelf;
  const int32_t* we;    // a batch's world per edge and per-world 1/n (else NULL: inv_n)
  const double* winv;
  const double *hedge, *acc;                  // air
  const double *hs, *us, *ubs, *bacc, *wall;  // ocean
};

__device__ __forceinline__ double stage_flux(const FluxSrc& f, const isw_grid& g, int l, int e) {
  const size_t o = (size_t)l * g.E + e;
  const double inv_n = f.we ? f.winv[f.we[e] * 3 + f.stage] : f.inv_n;
  if (f.air) return f.hedge[o] * f.acc[o] * inv_n;
  if (f.wall && f.wall[e] == 0) return 0.0;
  const double he = edge_h(f.hs[(size_t)l * g.N + g.eI[e]], f.hs[(size_t)l * g.N + g.eJ[e]], f.shelf);
  return he * (f.us[o] - f.ubs[e]) + he * (f.bacc[e] * inv_n);
}

__global__ __launch_bounds__(BS) void k_scale(isw_grid g, FluxSrc fs, double dts, double floor_,
                                              const double* __restrict__ h0, double* __restrict__ scale, int L) {
  const int t = XTID;
  if (t >= g.N * L) return;
  const int l = t / g.N, i = t - l * g.N;
  int e[6];
  ring6(g.cellEdgeAsc, i, e);
  double f[6];
  bool own[6];
#pragma unroll
  for (int m = 0; m < 6; m++) {
    const int ek = ok0(e[m]);
    f[m] = stage_flux(fs, g, l, ek) * g.le[ek] * dts;
    own[m] = g.eI[ek] == i;
  }
  double out = 0;
#pragma unroll
  for (int m = 0; m < 6; m++) {
    if (e[m] < 0) break;
    if (own[m]) { if (f[m] > 0) out += f[m]; }
    else if (!(f[m] > 0)) out -= f[m];
  }
  const double avail = h0[t] - floor_;
  const double cap = (avail > 0 ? avail : 0.0) * g.areaCell[i];
  scale[t] = out > cap ? cap / out : 1.0;
}

// lay_limit + lay_h: the limited fluxes summed into the new thickness; each edge's eI cell writes its Fm.
__global__ __launch_bounds__(BS) void k_limit_h(isw_grid g, Div dv, FluxSrc fs, double cdt, double dt,
                                                const double* __restrict__ scale, const double* h0, double* hn,
                                                double* __restrict__ Fm, int L) {
  const int t = XTID;
  if (t >= g.N * L) return;
  const int l = t / g.N, i = t - l * g.N;
  const double* sl = scale + (size_t)l * g.N;
  int e[6];
  ring6(g.cellEdge, i, e);
  double f[6], sc[6], le[6];
  int own[6];
#pragma unroll
  for (int m = 0; m < 6; m++) {
    const int ek = ok0(e[m]);
    f[m] = stage_flux(fs, g, l, ek);
    const int eI = g.eI[ek];
    own[m] = eI == i;
    sc[m] = f[m] > 0 ? sl[eI] : sl[g.eJ[ek]];
    le[m] = g.le[ek];
  }
  double div = 0;
#pragma unroll
  for (int m = 0; m < 6; m++) {
    if (e[m] < 0) break;
    const double fl = f[m] * sc[m];
    div += (double)g.cellSign[i * 6 + m] * le[m] * fl;
    if (Fm && own[m]) Fm[(size_t)l * g.E + e[m]] = fl * le[m] * dt;
  }
  const double v0 = h0[t];
  hn[t] = v0 - DIVBY(cdt * div, dv.A[i]);
}

__global__ __launch_bounds__(BS) void k_lay_u(int E, int L, double cdt, const double* u0, const double* __restrict__ Tp,
                                              const double* __restrict__ ub_end, const double* __restrict__ ub0,
                                              const double* __restrict__ wall, double* un) {
  const int t = XTID;
  if (t >= E * L) return;
  const int e = t % E;
  const double v = u0[t] + cdt * Tp[t] + (ub_end[e] - ub0[e]);
  un[t] = wall && wall[e] == 0 ? 0.0 : v;
}

// ---------------------------------------------------------------------------------------------------------------
// Interfacial stress, its heat, and the exchange between layers (layers.cu:454-621).
// ---------------------------------------------------------------------------------------------------------------
__global__ __launch_bounds__(BS) void k_lay_vmix(isw_grid g, int L, double c, double dt, double cd,
                                                 const double* __restrict__ h, const double* __restrict__ wall,
                                                 double* __restrict__ u, double* __restrict__ diss) {
  const int e = XTID;
  if (e >= g.E) return;
  const int N = g.N, E = g.E, eI = g.eI[e], eJ = g.eJ[e];
  double d = 0;
  for (int k = 0; k + 1 < L; k++) {
    const double ha = 0.5 * (h[(size_t)k * N + eI] + h[(size_t)k * N + eJ]);
    const double hb = 0.5 * (h[(size_t)(k + 1) * N + eI] + h[(size_t)(k + 1) * N + eJ]);
    double* ua = u + (size_t)k * E + e;
    double* ub = u + (size_t)(k + 1) * E + e;
    const double mean = (ha * *ua + hb * *ub) / (ha + hb);
    const double du0 = *ua - *ub;
    double du1 = du0 * exp(-c * (1 / ha + 1 / hb) * dt);
    if (cd > 0) du1 = du1 / (1 + cd * fabs(du0) * (1 / ha + 1 / hb) * dt);
    *ua = mean + hb / (ha + hb) * du1;
    *ub = mean - ha / (ha + hb) * du1;
    d += 0.5 * ha * hb / (ha + hb) * (du0 * du0 - du1 * du1);
  }
  diss[e] = d;
  if (wall && wall[e] == 0)
    for (int k = 0; k < L; k++) u[(size_t)k * E + e] = 0;
}

__global__ __launch_bounds__(BS) void k_vmix_heat(isw_grid g, LP p, const double* __restrict__ h,
                                                  const double* __restrict__ diss, double* __restrict__ th) {
  const int i = XTID;
  if (i >= g.N) return;
  const int N = g.N, L = p.L;
  int e[6];
  ring6(g.cellEdge, i, e);
  double t[6];
#pragma unroll
  for (int m = 0; m < 6; m++) t[m] = 0.5 * g.le[ok0(e[m])] * g.de[ok0(e[m])] * diss[ok0(e[m])];
  double q = 0;
#pragma unroll
  for (int m = 0; m < 6; m++)
    if (e[m] >= 0) q += t[m];
  if (q == 0) return;
  q /= g.areaCell[i];
  double mu = 0;
  for (int k = 0; k < L; k++) mu += h[(size_t)k * N + i];
  const double gr = grav_at(p, i);
  double pr = 0, xt = 0;
  for (int k = 0; k < L; k++) {
    const double pt = pr;
    pr = pr + gr * h[(size_t)k * N + i];
    const double x = pow(pr / p.p0, p.kappa);
    const double pib = p.cp * (pr * x - pt * xt) / ((p.kappa + 1) * (pr - pt));
    th[(size_t)k * N + i] += q / (mu * pib);
    xt = x;
  }
}

__global__ __launch_bounds__(BS) void k_xch_cell(int N, int L, int seq, const double* __restrict__ wx,
                                                 const double* __restrict__ areaCell, double* __restrict__ h,
                                                 double* __restrict__ M, double* __restrict__ psi) {
  const int i = XTID;
  if (i >= N) return;
  double tr[ISW_LMAX], dm[ISW_LMAX];
  for (int k = 0; k < L; k++) { tr[k] = 0; dm[k] = 0; }
  if (seq) {
    double m[ISW_LMAX], t[ISW_LMAX];
    for (int k = 0; k < L; k++) { const size_t o = (size_t)k * N + i; m[k] = M[o]; t[k] = M[o] * psi[o]; }
    for (int pass = 0; pass < 2; pass++)
      for (int n = 0; n + 1 < L; n++) {
        const int k = pass == 0 ? L - 2 - n : n;
        const double v = wx[(size_t)k * N + i];
        if (pass == 0 ? !(v > 0) : !(v < 0)) continue;
        const int d = v > 0 ? k + 1 : k;
        const double vol = v * areaCell[i], pp = t[d] / m[d];
        m[k] += vol; m[k + 1] -= vol; t[k] += vol * pp; t[k + 1] -= vol * pp;
      }
    for (int k = 0; k < L; k++) {
      const size_t o = (size_t)k * N + i;
      h[o] = h[o] + (m[k] - M[o]) / areaCell[i];
      M[o] = m[k];
      psi[o] = t[k] / m[k];
    }
    return;
  }
  for (int k = 0; k + 1 < L; k++) {
    const double v = wx[(size_t)k * N + i];
    const double pp = v > 0 ? psi[(size_t)(k + 1) * N + i] : psi[(size_t)k * N + i];
    const double vol = v * areaCell[i];
    dm[k] += vol; dm[k + 1] -= vol;
    tr[k] += vol * pp; tr[k + 1] -= vol * pp;
  }
  for (int k = 0; k < L; k++) {
    const size_t o = (size_t)k * N + i;
    const double m1 = M[o] + dm[k];
    psi[o] = (M[o] * psi[o] + tr[k]) / m1;
    M[o] = m1;
    h[o] = h[o] + dm[k] / areaCell[i];
  }
}

__global__ __launch_bounds__(BS) void k_xch_edge(isw_grid g, int L, const double* __restrict__ wx,
                                                 const double* __restrict__ hold, double* __restrict__ u,
                                                 double* __restrict__ diss, const double* __restrict__ lap4,
                                                 double nu4dt, double hthin, double floor_,
                                                 const double* __restrict__ wall) {
  const int e = XTID;
  if (e >= g.E) return;
  const int N = g.N, E = g.E, eI = g.eI[e], eJ = g.eJ[e];
  double mom[ISW_LMAX], hn[ISW_LMAX];
  double k0 = 0, v4 = 0;
  for (int k = 0; k < L; k++) {
    const double he = 0.5 * (hold[(size_t)k * N + eI] + hold[(size_t)k * N + eJ]);
    const double uk = u[(size_t)k * E + e];
    mom[k] = he * uk;
    hn[k] = he;
    k0 += he * uk * uk;
    if (lap4) v4 += he * uk * lap4[(size_t)k * E + e];
  }
  for (int pass = 0; pass < 2; pass++)
    for (int n = 0; n + 1 < L; n++) {
      const int k = pass == 0 ? L - 2 - n : n;
      const double v = 0.5 * (wx[(size_t)k * N + eI] + wx[(size_t)k * N + eJ]);
      if (pass == 0 ? !(v > 0) : !(v < 0)) continue;
      const int d = v > 0 ? k + 1 : k;
      const double ud = hn[d] > floor_ ? mom[d] / hn[d] : u[(size_t)d * E + e];
      mom[k] += v * ud; mom[k + 1] -= v * ud;
      hn[k] += v; hn[k + 1] -= v;
    }
  double un[ISW_LMAX];
  for (int k = 0; k < L; k++) un[k] = mom[k] / hn[k];
  if (hthin > 0)
    for (int k = 0; k < L; k++) {
      const double hm = fmin(hold[(size_t)k * N + eI], hold[(size_t)k * N + eJ]);
      if (hm >= hthin) continue;
      double w = (hthin - hm) / (hthin - floor_);
      if (w > 1) w = 1;
      double hs = 0, ms = 0;
      if (k > 0) { hs += hn[k - 1]; ms += hn[k - 1] * un[k - 1]; }
      if (k + 1 < L) { hs += hn[k + 1]; ms += hn[k + 1] * un[k + 1]; }
      if (hs <= 0) continue;
      w = fmin(w, hs / (hn[k] + hs));
      const double du = w * (ms / hs - un[k]);
      un[k] += du;
      const double dn = hn[k] * du / hs;
      if (k > 0) un[k - 1] -= dn;
      if (k + 1 < L) un[k + 1] -= dn;
    }
  double k1 = 0;
  const bool walled = wall && wall[e] == 0;
  for (int k = 0; k < L; k++) {
    u[(size_t)k * E + e] = walled ? 0.0 : un[k];
    k1 += hn[k] * un[k] * un[k];
  }
  if (diss) diss[e] = 0.5 * (k0 - k1) + nu4dt * v4;
}

// ---------------------------------------------------------------------------------------------------------------
// Transport (fused.cu:124-241; trisk.cu:48-59). Rows y = tracer * Lm + layer; a row's masses and fluxes are layer
// y % Lm's. blockIdx.y is the row.
// ---------------------------------------------------------------------------------------------------------------
__global__ __launch_bounds__(BS) void k_mass_update(isw_grid g, const double* __restrict__ M0,
                                                    const double* __restrict__ Fm, double* __restrict__ M1) {
  const int i = XTID;
  if (i >= g.N) return;
  const int l = blockIdx.y;
  const size_t o = (size_t)l * g.N + i;
  const double* f = Fm + (size_t)l * g.E;
  int e[6];
  ring6(g.cellEdgeAsc, i, e);
  double fv[6];
  bool own[6];
#pragma unroll
  for (int k = 0; k < 6; k++) { fv[k] = f[ok0(e[k])]; own[k] = g.eI[ok0(e[k])] == i; }
  double m = M0[o];
#pragma unroll
  for (int k = 0; k < 6; k++) {
    if (e[k] < 0) break;
    if (own[k]) m -= fv[k]; else m += fv[k];
  }
  M1[o] = m;
}

__device__ __forceinline__ double upwind(const double* Fml, const double* p, const isw_grid& g, int e) {
  const double f = Fml[e];
  return f * (f >= 0 ? p[g.eI[e]] : p[g.eJ[e]]);
}

// 3rd-order flux with upwind weight beta (fused.cu:131-138): f (mean - a - b) + |f| (b - a) beta; the curvature
// stencil on `psf` (the field itself in the plain form, its FP32 copy in the mixed form) with weights wt.
// The high-order flux of one edge from its stencil (eight cells ix with weights w, its two cells I, J), for the
// tracer row ps (psf its stencil copy) and the edge's mass flux f.
__device__ __forceinline__ double hoflux_at(double f, const double* ps, const fast* psf, const int (&ix)[8],
                                            const fast (&w)[8], int I, int J, double beta) {
  const fast a = w[0] * psf[ix[0]] + w[1] * psf[ix[1]] + w[2] * psf[ix[2]] + w[3] * psf[ix[3]];
  const fast b = w[4] * psf[ix[4]] + w[5] * psf[ix[5]] + w[6] * psf[ix[6]] + w[7] * psf[ix[7]];
  return f * (0.5 * (ps[I] + ps[J]) - (double)(a + b)) + fabs(f) * ((double)(b - a) * beta);
}
__device__ __forceinline__ void stencil_at(const isw_grid& g, const fast* wt, int e, int (&ix)[8], fast (&w)[8]) {
  const int o = e * 8;
#pragma unroll
  for (int k = 0; k < 8; k++) {
    ix[k] = g.d2idx[o + k];
    w[k] = wt[o + k];
  }
}
__device__ __forceinline__ double hoflux(const double* Fml, const double* ps, const fast* psf, const fast* wt,
                                         const isw_grid& g, int e, double beta) {
  int ix[8];
  fast w[8];
  stencil_at(g, wt, e, ix, w);
  return hoflux_at(Fml[e], ps, psf, ix, w, g.eI[e], g.eJ[e], beta);
}

__global__ __launch_bounds__(BS) void k_tr_td(isw_grid g, int Lm, int rows, const double* __restrict__ M0,
                                              const double* __restrict__ psi, const double* __restrict__ Fm,
                                              const double* __restrict__ M1, double* __restrict__ td,
                                              float* __restrict__ pf) {
  const int i = XTID;
  const int y = blockIdx.y * blockDim.y + threadIdx.y, l = y % Lm;
  if (i >= g.N || y >= rows) return;
  const size_t o = (size_t)y * g.N + i, om = (size_t)l * g.N + i;
  const double* p = psi + (size_t)y * g.N;
  const double* Fml = Fm + (size_t)l * g.E;
  int e[6];
  ring6(g.cellEdgeAsc, i, e);
  double fv[6];
  bool own[6];
#pragma unroll
  for (int k = 0; k < 6; k++) { fv[k] = upwind(Fml, p, g, ok0(e[k])); own[k] = g.eI[ok0(e[k])] == i; }
  double t = M0[om] * psi[o];
#pragma unroll
  for (int k = 0; k < 6; k++) {
    if (e[k] < 0) break;
    if (own[k]) t -= fv[k]; else t += fv[k];
  }
  td[o] = t / M1[om];
  if (pf) pf[o] = (float)psi[o];
}

__global__ __launch_bounds__(BS) void k_tr_hflux(isw_grid g, int Lm, int rows, double frac, double beta,
                                                 const double* __restrict__ src, const fast* __restrict__ srcf,
                                                 const fast* __restrict__ wt, const double* __restrict__ Fm,
                                                 double* __restrict__ Hf) {
  const int e = XTID;
  const int y = blockIdx.y * blockDim.y + threadIdx.y, l = y % Lm;
  if (e >= g.E || y >= rows) return;
  Hf[(size_t)y * g.E + e] =
      frac * hoflux(Fm + (size_t)l * g.E, src + (size_t)y * g.N, srcf + (size_t)y * g.N, wt, g, e, beta);
}

__global__ __launch_bounds__(BS) void k_tr_stage(isw_grid g, int Lm, int rows, double frac, const double* __restrict__ M0,
                                                 const double* __restrict__ psi, const double* __restrict__ Hf,
                                                 const double* __restrict__ Fm, double* __restrict__ out,
                                                 float* __restrict__ pf) {
  const int i = XTID;
  const int y = blockIdx.y * blockDim.y + threadIdx.y, l = y % Lm;
  if (i >= g.N || y >= rows) return;
  const size_t o = (size_t)y * g.N + i, om = (size_t)l * g.N + i;
  const double* Hl = Hf + (size_t)y * g.E;
  const double* Fml = Fm + (size_t)l * g.E;
  int e[6];
  ring6(g.cellEdgeAsc, i, e);
  double fv[6], mv[6];
  bool own[6];
#pragma unroll
  for (int k = 0; k < 6; k++) {
    const int ek = ok0(e[k]);
    fv[k] = Hl[ek];
    mv[k] = frac * Fml[ek];
    own[k] = g.eI[ek] == i;
  }
  double t = M0[om] * psi[o];
  double mx = M0[om];
#pragma unroll
  for (int k = 0; k < 6; k++) {
    if (e[k] < 0) break;
    if (own[k]) { t -= fv[k]; mx -= mv[k]; } else { t += fv[k]; mx += mv[k]; }
  }
  const double r = t / mx;
  out[o] = r;
  if (pf) pf[o] = (float)r;
}

// One RK stage of the transport in one launch on the grid's 64-cell patches (k_tr_hflux then k_tr_stage): block
// (patch, four rows) first computes the third-order flux of every edge of its cells -- its own edges and those
// whose second cell only is in the patch, the latter also computed by the neighbour that owns them, from the same
// inputs, so every copy is the same to the bit -- into LDS, then steps its 64 cells x 4 rows from there. The flux
// array's write and re-read (rows x E doubles each way) and one kernel boundary go away; the arithmetic and every
// sum's order are k_tr_hflux's and k_tr_stage's. pfo: the output's FP32 copy (mixed form; the stencil reads srcf,
// a different buffer, so no block overwrites what another still reads).
__global__ __launch_bounds__(BS) void k_tr_hstage(isw_grid g, int Lm, int rows, double frac, double beta,
                                                  const double* __restrict__ src, const fast* __restrict__ srcf,
                                                  const fast* __restrict__ wt, const double* __restrict__ Fm,
                                                  const double* __restrict__ M0, const double* __restrict__ psi,
                                                  double* __restrict__ out, float* __restrict__ pfo) {
  extern __shared__ double hs[];                   // [4][ptedges] the stage flux of the patch's edges, per row
  const int PT = g.ptedges, N = g.N, E = g.E;
  const int b = xblock();
  const int tx = threadIdx.x, ty = threadIdx.y;
  const int y = blockIdx.y * 4 + ty;
  const bool rowok = y < rows;
  const int l = y % Lm;
  const int e0 = g.pe_off[b], ne = g.pe_off[b + 1] - e0;
  const int i0 = g.pi_off[b], ni = g.pi_off[b + 1] - i0;
  // the stage flux of the patch's edges for the block's four rows: thread t takes edges t, t + BS, ..., each edge's
  // stencil (its eight cells and weights, its two cells) loaded once for the four rows
  const int y0 = blockIdx.y * 4, nq = ne + ni;
  for (int q = ty * BS1 + tx; q < nq; q += BS) {
    const int e = q < ne ? g.pe_edge[e0 + q] : g.pi_edge[i0 + q - ne];
    int ix[8];
    fast w[8];
    stencil_at(g, wt, e, ix, w);
    const int I = g.eI[e], J = g.eJ[e];
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int yr = y0 + r;
      if (yr < rows) {
        const double f = Fm[(size_t)(yr % Lm) * E + e];
        hs[r * PT + q] = frac * hoflux_at(f, src + (size_t)yr * N, srcf + (size_t)yr * N, ix, w, I, J, beta);
      }
    }
  }
  __syncthreads();
  const int i = b * BS1 + tx;
  if (!rowok || i >= N) return;
  const size_t o = (size_t)y * N + i, om = (size_t)l * N + i;
  const double* Fml = Fm + (size_t)l * E;
  int e[6], sl[6];
  ring6(g.cellEdgeAsc, i, e);
  ring6(g.tc_slot, i, sl);
  double fv[6], mv[6];
#pragma unroll
  for (int k = 0; k < 6; k++) {
    fv[k] = hs[ty * PT + (sl[k] < 0 ? 0 : sl[k] >> 1)];
    mv[k] = frac * Fml[ok0(e[k])];
  }
  double t = M0[om] * psi[o];
  double mx = M0[om];
#pragma unroll
  for (int k = 0; k < 6; k++) {
    if (e[k] < 0) break;
    if (sl[k] & 1) { t -= fv[k]; mx -= mv[k]; } else { t += fv[k]; mx += mv[k]; }
  }
  const double r = t / mx;
  out[o] = r;
  if (pfo) pfo[o] = (float)r;
}

__global__ __launch_bounds__(BS) void k_tr_aflux(isw_grid g, int Lm, int rows, double beta, const double* __restrict__ psi,
                                                 const double* __restrict__ p2, const fast* __restrict__ p2f,
                                                 const fast* __restrict__ wt, const double* __restrict__ Fm,
                                                 double* __restrict__ A) {
  // one edge per thread (256 per block), its stencil loaded once for the block's four rows
  const int e = xblock() * BS + threadIdx.y * BS1 + threadIdx.x;
  if (e >= g.E) return;
  int ix[8];
  fast w[8];
  stencil_at(g, wt, e, ix, w);
  const int I = g.eI[e], J = g.eJ[e];
  const int y0 = blockIdx.y * 4;
#pragma unroll
  for (int r = 0; r < 4; r++) {
    const int y = y0 + r;
    if (y < rows) {
      const double f = Fm[(size_t)(y % Lm) * g.E + e];
      const double* p = psi + (size_t)y * g.N;
      A[(size_t)y * g.E + e] = hoflux_at(f, p2 + (size_t)y * g.N, p2f + (size_t)y * g.N, ix, w, I, J, beta) -
                               f * (f >= 0 ? p[I] : p[J]);
    }
  }
}

__global__ __launch_bounds__(BS) void k_tr_ratios(isw_grid g, int Lm, int rows, const double* __restrict__ psi,
                                                  const double* __restrict__ td, const double* __restrict__ M1,
                                                  const double* __restrict__ A, double* __restrict__ Rp,
                                                  double* __restrict__ Rm) {
  const int i = XTID;
  const int y = blockIdx.y * blockDim.y + threadIdx.y, l = y % Lm;
  if (i >= g.N || y >= rows) return;
  const double* p = psi + (size_t)y * g.N;
  const double* t = td + (size_t)y * g.N;
  const double* Al = A + (size_t)y * g.E;
  int nb[6], e[6];
  ring6(g.nbr, i, nb);
  ring6(g.cellEdgeAsc, i, e);
  double pv[6], tv[6], av[6];
  bool own[6];
#pragma unroll
  for (int k = 0; k < 6; k++) {
    pv[k] = p[ok0(nb[k])];
    tv[k] = t[ok0(nb[k])];
    av[k] = Al[ok0(e[k])];
    own[k] = g.eI[ok0(e[k])] == i;
  }
  double hi = fmax(p[i], t[i]), lo = fmin(p[i], t[i]);
#pragma unroll
  for (int k = 0; k < 6; k++) {
    if (nb[k] < 0) break;
    hi = fmax(hi, fmax(pv[k], tv[k]));
    lo = fmin(lo, fmin(pv[k], tv[k]));
  }
  const double m1 = M1[(size_t)l * g.N + i];
  const double qmax = (hi - t[i]) * m1, qmin = (t[i] - lo) * m1;
  double Pp = 0, Pm = 0;
#pragma unroll
  for (int k = 0; k < 6; k++) {
    if (e[k] < 0) break;
    const double a = av[k];
    if (own[k]) { if (a > 0) Pm += a; else Pp -= a; }
    else { if (a > 0) Pp += a; else Pm -= a; }
  }
  Rp[(size_t)y * g.N + i] = Pp > 0 ? fmin(1.0, qmax / Pp) : 0.0;
  Rm[(size_t)y * g.N + i] = Pm > 0 ? fmin(1.0, qmin / Pm) : 0.0;
}

__global__ __launch_bounds__(BS) void k_tr_apply(isw_grid g, int Lm, int rows, const double* __restrict__ td,
                                                 const double* __restrict__ M1, const double* __restrict__ A,
                                                 const double* __restrict__ Rp, const double* __restrict__ Rm,
                                                 double* __restrict__ out) {
  const int i = XTID;
  const int y = blockIdx.y * blockDim.y + threadIdx.y, l = y % Lm;
  if (i >= g.N || y >= rows) return;
  const size_t o = (size_t)y * g.N + i, om = (size_t)l * g.N + i;
  const double* Al = A + (size_t)y * g.E;
  const double* rp = Rp + (size_t)y * g.N;
  const double* rm = Rm + (size_t)y * g.N;
  int e[6];
  ring6(g.cellEdgeAsc, i, e);
  double av[6], cv[6];
  bool own[6];
#pragma unroll
  for (int k = 0; k < 6; k++) {
    const int ek = ok0(e[k]);
    const int eI = g.eI[ek], eJ = g.eJ[ek];
    av[k] = Al[ek];
    cv[k] = av[k] > 0 ? fmin(rp[eJ], rm[eI]) : fmin(rp[eI], rm[eJ]);
    own[k] = eI == i;
  }
  double t = td[o] * M1[om];
#pragma unroll
  for (int k = 0; k < 6; k++) {
    if (e[k] < 0) break;
    if (own[k]) t -= cv[k] * av[k]; else t += cv[k] * av[k];
  }
  out[o] = t / M1[om];
}

// ---------------------------------------------------------------------------------------------------------------
// SWE (fused.cu:27-120): five passes per stage.
// ---------------------------------------------------------------------------------------------------------------
__global__ __launch_bounds__(BS) void k_swe_scale(isw_grid g, double dts, const double* __restrict__ h,
                                                  const double* __restrict__ u, double* __restrict__ scale) {
  const int i = XTID;
  if (i >= g.N) return;
  const int l = blockIdx.y;
  const double* hl = h + (size_t)l * g.N;
  const double* ul = u + (size_t)l * g.E;
  double out = 0;
#pragma unroll
  for (int k = 0; k < 6; k++) {
    const int e = g.cellEdgeAsc[i * 6 + k];
    if (e < 0) break;
    const double f = 0.5 * (hl[g.eI[e]] + hl[g.eJ[e]]) * ul[e] * g.le[e] * dts;
    if (g.eI[e] == i) { if (f > 0) out += f; }
    else if (!(f > 0)) out -= f;
  }
  const double hv = hl[i];
  const double cap = 0.5 * (hv > 0 ? hv : 0.0) * g.areaCell[i];
  scale[(size_t)l * g.N + i] = out > cap ? cap / out : 1.0;
}

__global__ __launch_bounds__(BS) void k_swe_flux(isw_grid g, double dt, const double* __restrict__ h,
                                                 const double* __restrict__ u, const double* __restrict__ scale,
                                                 double* __restrict__ F, double* __restrict__ Fm) {
  const int e = XTID;
  if (e >= g.E) return;
  const int l = blockIdx.y;
  const size_t o = (size_t)l * g.E + e;
  const double* hl = h + (size_t)l * g.N;
  double f = 0.5 * (hl[g.eI[e]] + hl[g.eJ[e]]) * u[o];
  f *= f > 0 ? scale[(size_t)l * g.N + g.eI[e]] : scale[(size_t)l * g.N + g.eJ[e]];
  F[o] = f;
  if (Fm) Fm[o] = f * g.le[e] * dt;
}

__global__ __launch_bounds__(BS) void k_swe_cell(isw_grid g, double grav, double c, const double* __restrict__ h,
                                                 const double* __restrict__ b, const double* __restrict__ F,
                                                 const double* __restrict__ u, const double* h0, double* hn,
                                                 double* __restrict__ B) {
  const int i = XTID;
  if (i >= g.N) return;
  const int l = blockIdx.y;
  const double* Fl = F + (size_t)l * g.E;
  const double* ul = u + (size_t)l * g.E;
  double div = 0, ke = 0;
#pragma unroll
  for (int k = 0; k < 6; k++) {
    const int e = g.cellEdge[i * 6 + k];
    if (e < 0) break;
    div += (double)g.cellSign[i * 6 + k] * g.le[e] * Fl[e];
    ke += g.le[e] * g.de[e] * 0.25 * ul[e] * ul[e];
  }
  const size_t o = (size_t)l * g.N + i;
  const double dh = 0.0 - div / g.areaCell[i];
  B[o] = grav * (h[o] + b[i]) + 0.0 + ke / g.areaCell[i];
  hn[o] = h0[o] + c * dh;
}

__global__ __launch_bounds__(BS) void k_swe_vertex(isw_grid g, Div dv, const double* __restrict__ h,
                                                   const double* __restrict__ u, const double* __restrict__ fV,
                                                   double* __restrict__ q) {
  const int v = XTID;
  if (v >= g.NV) return;
  const int l = blockIdx.y;
  const double* hl = h + (size_t)l * g.N;
  const double* ul = u + (size_t)l * g.E;
  double hv = 0, z = 0;
#pragma unroll
  for (int k = 0; k < 3; k++) {
    const int c = g.vCell[v * 3 + k];
    hv += g.kite[c * 6 + g.vSlot[v * 3 + k]] * hl[c] * 1.0;
  }
#pragma unroll
  for (int k = 0; k < 3; k++) {
    const int e = g.vEdge[v * 3 + k];
    z += (double)g.vEdgeSign[v * 3 + k] * g.de[e] * ul[e];
  }
  q[(size_t)l * g.NV + v] = (DIVBY(z, dv.AV[v]) + fV[v]) / DIVBY(hv, dv.AV[v]);
}

// ---------------------------------------------------------------------------------------------------------------
// Air-sea stress, the land column, the mixed layer's capacity (ocean.cu:17-42,434-471).
// ---------------------------------------------------------------------------------------------------------------
__global__ __launch_bounds__(BS) void k_couple_stress(isw_grid g, int La, double c0, double rho_w, double dt,
                                                      double land_drag0, const double* __restrict__ ha,
                                                      double* __restrict__ ua, const double* __restrict__ ho,
                                                      double* __restrict__ uo, double* __restrict__ diss,
                                                      double* __restrict__ tau_acc, const double* __restrict__ wall,
                                                      const double* __restrict__ cd) {
  const int e = XTID;
  if (e >= g.E) return;
  const int N = g.N, E = g.E, b = La - 1, eI = g.eI[e], eJ = g.eJ[e];
  double c = c0, land_drag = land_drag0;
  if (cd) { c = 0.5 * (cd[eI] + cd[eJ]); land_drag = 1; }
  const double ma = 0.5 * (ha[(size_t)b * N + eI] + ha[(size_t)b * N + eJ]);
  const size_t oa = (size_t)b * E + e;
  const double a0 = ua[oa];
  double a1;
  if (wall && wall[e] == 0) {
    const double du1 = a0 * exp(-land_drag * c / ma * dt);
    a1 = du1;
    diss[e] = 0.5 * ma * (a0 * a0 - du1 * du1);
  } else {
    const double o0 = uo[e];
    const double mo = rho_w * 0.5 * (ho[eI] + ho[eJ]);
    const double mean = (ma * a0 + mo * o0) / (ma + mo);
    const double du0 = a0 - o0;
    const double du1 = du0 * exp(-c * (1 / ma + 1 / mo) * dt);
    a1 = mean + mo / (ma + mo) * du1;
    uo[e] = mean - ma / (ma + mo) * du1;
    diss[e] = 0.5 * ma * mo / (ma + mo) * (du0 * du0 - du1 * du1);
  }
  ua[oa] = a1;
  const double tau = ma * (a0 - a1) / dt;
  tau_acc[e] += tau * tau * dt;
}

__global__ __launch_bounds__(BS) void k_land_column(int N, int K, double c_soil, double k_soil, double dt,
                                                    const double* __restrict__ land, double* __restrict__ T,
                                                    const double* __restrict__ cap, double* __restrict__ Tg,
                                                    const double* __restrict__ dz) {
  const int i = XTID;
  if (i >= N) return;
  if (!(land[i] > 0.5)) return;
  const int n = K + 1;
  double C[ISW_KSOIL + 1], G[ISW_KSOIL + 1], x[ISW_KSOIL + 1], cp_[ISW_KSOIL + 1], dp[ISW_KSOIL + 1];
  C[0] = cap[i];
  x[0] = T[i];
  double half = cap[i] / c_soil / 2;
  for (int k = 1; k < n; k++) {
    const double d = dz[k - 1];
    C[k] = c_soil * d;
    x[k] = Tg[(size_t)(k - 1) * N + i];
    G[k - 1] = dt * k_soil / (half + d / 2);
    half = d / 2;
  }
  G[n - 1] = 0;
  for (int k = 0; k < n; k++) {
    const double a = k ? G[k - 1] : 0.0, c = G[k];
    const double den = C[k] + a + c - (k ? a * cp_[k - 1] : 0.0);
    cp_[k] = c / den;
    dp[k] = (C[k] * x[k] + (k ? a * dp[k - 1] : 0.0)) / den;
  }
  x[n - 1] = dp[n - 1];
  for (int k = n - 2; k >= 0; k--) x[k] = dp[k] + cp_[k] * x[k + 1];
  T[i] = x[0];
  for (int k = 1; k < n; k++) Tg[(size_t)(k - 1) * N + i] = x[k];
}

__global__ __launch_bounds__(BS) void k_ocn_cap(int N, double rho_cw, double c_land, const double* __restrict__ h,
                                                double* __restrict__ cap, const double* __restrict__ land) {
  const int i = XTID;
  if (i >= N) return;
  cap[i] = land && land[i] > 0.5 ? (cap[i] > 0 ? cap[i] : c_land) : rho_cw * h[i];
}

// SWE's edge pass (swe_edge, fused.cu:101-120) with each layer's drag from the layer settings passed by value.
__global__ __launch_bounds__(BS) void k_swe_edge_lp(isw_grid g, LP p, double c, const double* __restrict__ F,
                                                    const double* __restrict__ q, const double* __restrict__ B,
                                                    const double* __restrict__ u, const double* __restrict__ lap4,
                                                    const double* u0, double* un) {
  const int e = XTID;
  if (e >= g.E) return;
  const int l = blockIdx.y;
  const double* Fl = F + (size_t)l * g.E;
  const double* ql = q + (size_t)l * g.NV;
  const double* Bl = B + (size_t)l * g.N;
  const double qa = 0.5 * (ql[g.ev1[e]] + ql[g.ev2[e]]);
  double Q = 0;
  for (int k = g.eoeStart[e]; k < g.eoeStart[e + 1]; k++) {
    const int e2 = g.eoe[k];
    const double qb = 0.5 * (ql[g.ev1[e2]] + ql[g.ev2[e2]]);
    Q += g.w[k] * Fl[e2] * 0.5 * (qa + qb);
  }
  const size_t o = (size_t)l * g.E + e;
  double d = Q - (Bl[g.eJ[e]] - Bl[g.eI[e]]) / g.de[e] - p.drag[l] * u[o];
  if (p.nu4 != 0) d = d - p.nu4 * lap4[o];
  un[o] = u0[o] + c * d;
}

// ===============================================================================================================
// Host sequencing (the reference's Python drivers), all on one stream. Every launch goes through ISW_LAUNCH with the
// kernel's algorithmic bytes for the measurement mode (diag.hpp).
// ===============================================================================================================

// Byte counts: D doubles, I int32, C int8, R fast (the PV-flux and stencil operands' width).
constexpr double D = 8, I = 4, C = 1, R = sizeof(fast);

static double flux_bytes(const FluxSrc& f, double N, double E, double L) {
  return f.air ? 2 * L * E * D : L * N * D + L * E * D + 2 * E * D + (f.wall ? E * D : 0);
}

static void lap(const isw_grid* g, int L, const double* u, double* D1, double* Z1, double* out, hipStream_t st) {
  const double N = g->N, E = g->E, NV = g->NV, l = L;
  const int nbc = (g->N * L + BS - 1) / BS, nbv = (g->NV * L + BS - 1) / BS;
  ISW_LAUNCH("div_curl", l * E * D + 6 * N * (I + C) + E * D + N * D + l * N * D + 3 * NV * (I + C) + E * D + NV * D +
                             l * NV * D,
             k_div_curl, dim3(nbc + nbv), BS, *g, L, nbc, u, D1, Z1);
  ISW_LAUNCH("edge_lap", l * N * D + l * NV * D + 4 * E * I + 2 * E * D + l * E * D,
             k_edge_lap, g1((long)g->E * L), BS, *g, L, D1, Z1, out);
}

// Layers._stage (layers.py:255-309).
static void stage(const isw_grid* g, const isw_layers* s, const LP& p, const Div& dv, int si, const double* hs,
                  const double* us, const double* h0, const double* u0, double c, double dt, int nfast, double* hn,
                  double* un, double* Fm, const double* theta, hipStream_t st) {
  const int N = g->N, E = g->E, L = s->L;
  const double n_ = N, e_ = E, v_ = g->NV, l_ = L, air = s->air ? 1 : 0;
  const int nbc = (N + BS1 - 1) / BS1, nbv = (g->NV + BS1 - 1) / BS1;
  ISW_LAUNCH("bern_pv", l_ * n_ * D * (1 + 3 * air) + l_ * e_ * D + n_ * D + 6 * n_ * I + 2 * e_ * D + n_ * D +
                            l_ * n_ * D + 3 * v_ * (2 * I + 2 * C) + 6 * n_ * D + 2 * v_ * D + l_ * v_ * D,
             k_bern_pv, dim3(nbc + nbv), dim3(BS1, L), *g, p, dv, nbc, hs, us, theta, s->b, s->S, s->fV, s->B,
                                                      s->pibar, s->q);
#if ISW_FORM_MIXED
  fast* pk = s->pkf;
  const fast* wf = g->wf;
  // the air's fast sub-steps' coefficients per cell slot (FP32 in the mixed form)
  fast* cf = s->air && g->ph_off ? s->cfcf : nullptr;
#else
  fast* pk = s->pk;
  const fast* wf = g->w;
  fast* cf = s->air && g->ph_off ? s->cfc : nullptr;
#endif
  ISW_LAUNCH("edge_hq", l_ * n_ * D + l_ * v_ * D + l_ * e_ * D + 4 * e_ * I + l_ * e_ * D + 3 * l_ * e_ * R +
                            (cf ? 6 * l_ * n_ * sizeof(fast) + 2 * e_ + 6 * n_ + e_ * D : 0),
             k_edge_hq, g1((long)E * L), BS, *g, L, s->shelf, hs, s->q, us, s->hedge, pk, cf);
  const double neoe = g->n_eoe;
  const double slow_bytes = 3 * l_ * e_ * R + neoe * (R + I) + (e_ + 1) * I + 2 * l_ * e_ * D + l_ * n_ * D +
                            2 * e_ * I + e_ * D + (s->nu4 != 0 ? l_ * e_ * D : 0) + air * 2 * l_ * n_ * D + l_ * e_ * D +
                            3 * e_ * D;
#define SLOW(LL)                                                                                                   \
  ISW_LAUNCH("slow", slow_bytes, k_slow_e<LL>, g1(E), BS, *g, p, dv, c * dt, pk, wf, s->hedge, us, s->B,             \
             s->nu4 != 0 ? s->L2 : nullptr, s->air ? theta : nullptr, s->air ? s->pibar : nullptr, s->Tp, s->G, s->He, \
             s->ubs)
  switch (L) {
    case 1: SLOW(1); break;
    case 2: SLOW(2); break;
    case 3: SLOW(3); break;
    case 4: SLOW(4); break;
    case 5: SLOW(5); break;
    case 6: SLOW(6); break;
    case 7: SLOW(7); break;
    default: SLOW(8); break;
  }
#undef SLOW
  int n = (int)std::ceil(c * nfast);
  if (n < 1) n = 1;
  const double tau = c * dt / n;
  FluxSrc fs{};
  fs.inv_n = 1.0 / n;
  fs.shelf = s->shelf;
  fs.stage = si;
  if (s->batch) {                                  // every world's own sub-steps: the loop runs the largest count
    n = s->batch->nmax[si];
    fs.we = s->batch->we;
    fs.winv = s->batch->inv_n;
  }
  const size_t patch_lds = g->ph_off ? (size_t)L * g->pslots * sizeof(double) : 0;
  if (s->air && g->ph_off && patch_lds <= 65536) {
    // one launch per sub-step on the grid's patches; masses and winds alternate between (muf, mup) and (uf, ufb)
    // the last stage steps the wind in place (un is u0): where a launch might both read u0 (its first sub-step's
    // halo) and write un (an element's last sub-step), the sub-steps read a copy of u0 instead (s->F is free here)
    const double* uin = u0;
    if (un == u0 && (s->batch || n <= 2)) {
      ISW_LAUNCH("copy", 2 * l_ * e_ * D, k_copy, g1((long)E * L), BS, (long)E * L, u0, s->F);
      uin = s->F;
    }
    const double* mus = h0;
    // LDS of the pair kernel: Phi* twice, masses and winds; with L <= 4 also its patch-and-halo cells' coefficients
    const size_t pair_lds =
        g->ph2_off ? (size_t)L * (3 * g->pslots2 + g->pedges2 + (L <= 4 ? 6 * g->pslots : 0)) * sizeof(double) : 0;
    int it = 0;
    int flip = 0;
    if (g->ph2_off && pair_lds <= 65536) {
      // two sub-steps per launch while two remain; an odd one left runs alone below
      for (; it + 1 < n; it += 2, flip ^= 1) {
        double* uo = flip ? s->ufb : s->uf;
        double* mo = flip ? s->mup : s->muf;
        const double bytes = l_ * e_ * D + 6 * l_ * n_ * sizeof(fast) + 3 * l_ * n_ * D + l_ * e_ * D + e_ * D +
                             (it == 0 ? 1 : 2) * l_ * e_ * D + l_ * e_ * D + 6 * n_ * I + n_ * D + e_ * D +
                             2 * e_ * I;
#define PAIR(LL)                                                                                                  \
  ISW_LAUNCH_SHM("af_patch2", bytes, k_af_patch2<LL>, dim3((unsigned)g->npatch), dim3(FBS), pair_lds, *g, p, dv,  \
                 si, it, n, tau, cf, uin, mus, mo, s->S, s->Tp, s->G, s->accL, uo, un)
        switch (L) {
          case 1: PAIR(1); break;
          case 2: PAIR(2); break;
          case 3: PAIR(3); break;
          case 4: PAIR(4); break;
          case 5: PAIR(5); break;
          case 6: PAIR(6); break;
          case 7: PAIR(7); break;
          default: PAIR(8); break;
        }
#undef PAIR
        uin = uo;
        mus = mo;
      }
    }
    for (; it < n; it++, flip ^= 1) {
      double* uo = flip ? s->ufb : s->uf;
      double* mo = flip ? s->mup : s->muf;
      // (algorithmic bytes: uin, the cells' slot coefficients, mu in and out, S, Tp, G, acc, the wind out, the
      // cell ring, 1/A, de, the patch's edge list and slots; the halo's re-reads and Phi* in LDS are free)
      const double bytes = l_ * e_ * D + 6 * l_ * n_ * sizeof(fast) + 3 * l_ * n_ * D + l_ * e_ * D + e_ * D +
                           (it == 0 ? 1 : 2) * l_ * e_ * D + l_ * e_ * D + 6 * n_ * I + n_ * D + e_ * D +
                           2 * e_ * I;
#define PATCH(LL)                                                                                                 \
  ISW_LAUNCH_SHM("af_patch", bytes, k_af_patch<LL>, dim3((unsigned)g->npatch), dim3(FBS), patch_lds, *g, p, dv, si, \
                 it, n, tau, cf, uin, mus, mo, s->S, s->Tp, s->G, s->accL, uo, un)
      switch (L) {
        case 1: PATCH(1); break;
        case 2: PATCH(2); break;
        case 3: PATCH(3); break;
        case 4: PATCH(4); break;
        case 5: PATCH(5); break;
        case 6: PATCH(6); break;
        case 7: PATCH(7); break;
        default: PATCH(8); break;
      }
#undef PATCH
      uin = uo;
      mus = mo;
    }
    fs.air = 1;
    fs.hedge = s->hedge;
    fs.acc = s->accL;
  } else if (s->air) {
    for (int it = 0; it < n; it++) {
      const bool first = it == 0;
      ISW_LAUNCH("af_muphi", 2 * l_ * e_ * D + 6 * n_ * (I + C) + e_ * D + n_ * D + 4 * l_ * n_ * D,
                 k_af_muphi, c1(N), dim3(BS1, L), *g, p, dv, si, it, n, tau, s->hedge, first ? u0 : s->uf,
                                                             first ? h0 : s->muf, s->muf, s->S, s->phi);
      ISW_LAUNCH("af_u", l_ * n_ * D + l_ * e_ * D + e_ * D + 2 * e_ * I + e_ * D + 2 * l_ * e_ * D +
                             (first ? 1 : 2) * l_ * e_ * D,
                 k_af_u, g1((long)E * L), BS, *g, L, dv, p, si, it, n, tau, s->phi, s->Tp, s->G,
                 first ? u0 : s->uf, s->accL, s->uf, un);
    }
    fs.air = 1;
    fs.hedge = s->hedge;
    fs.acc = s->accL;
  } else {
    if (g->ph_off && s->ubc) {
      // one launch per sub-step on the grid's patches: heights alternate between eta and eta_prev, winds between
      // ubb and ubc, each edge's last sub-step writes ub
      const double* ubs_ = s->ub0;
      const double* ets = s->eta0;
      for (int it = 0; it < n; it++) {
        double* uo = (it & 1) ? s->ubc : s->ubb;
        double* eo = (it & 1) ? s->eta_prev : s->eta;
        ISW_LAUNCH_SHM("bt_patch", 2 * e_ * D + 6 * n_ * (I + C) + e_ * D + n_ * D + 2 * n_ * D + e_ * D + e_ * D +
                                       n_ * D + (s->wall ? e_ * D : 0) + e_ * D + (it == 0 ? 1 : 2) * e_ * D +
                                       2 * e_ * I,
                       k_bt_patch, dim3((unsigned)g->npatch), dim3(FBS), (size_t)g->pslots * sizeof(double), *g, dv,
                       p, si, it, n, tau, s->div_damp, s->He, ubs_, ets, eo, s->G, s->Gc, s->wall, s->acc, uo,
                       s->ub);
        ubs_ = uo;
        ets = eo;
      }
    } else
    for (int it = 0; it < n; it++) {
      const bool first = it == 0;
      ISW_LAUNCH("bt_eta", 2 * e_ * D + 6 * n_ * (I + C) + e_ * D + n_ * D + 3 * n_ * D,
                 k_bt_eta, c1(N), BS1, *g, dv, p, si, it, n, tau, s->He, first ? s->ub0 : s->ub,
                                                 first ? s->eta0 : s->eta, s->eta, s->eta_prev);
      ISW_LAUNCH("bt_u", 3 * n_ * D + e_ * D + 2 * e_ * I + e_ * D + (s->wall ? e_ * D : 0) + 2 * e_ * D +
                             (first ? 1 : 2) * e_ * D,
                 k_bt_u, g1(E), BS, *g, dv, p, si, it, n, tau, s->div_damp, s->eta, s->eta_prev, s->G,
                                              s->Gc, s->wall, first ? s->ub0 : s->ub, s->ub, s->acc);
    }
    fs.air = 0;
    fs.hs = hs;
    fs.us = us;
    fs.ubs = s->ubs;
    fs.bacc = s->acc;
    fs.wall = s->wall;
  }
  const double fb = flux_bytes(fs, n_, e_, l_);
  ISW_LAUNCH("scale", fb + 6 * n_ * I + 2 * e_ * I + e_ * D + l_ * n_ * D + n_ * D + l_ * n_ * D,
             k_scale, g1((long)N * L), BS, *g, fs, c * dt, s->floor_, h0, s->scale, L);
  ISW_LAUNCH("limit_h", fb + 6 * n_ * (I + C) + 2 * e_ * I + e_ * D + 3 * l_ * n_ * D + n_ * D +
                            (Fm ? l_ * e_ * D : 0),
             k_limit_h, g1((long)N * L), BS, *g, dv, fs, c * dt, dt, s->scale, h0, hn, Fm, L);
  if (!s->air)
    ISW_LAUNCH("lay_u", 3 * l_ * e_ * D + 2 * e_ * D + (s->wall ? e_ * D : 0),
               k_lay_u, g1((long)E * L), BS, E, L, c * dt, u0, s->Tp, s->ub, s->ub0, s->wall, un);
}

int layers_step(const isw_grid* g, const isw_layers* s, double* h, double* u, double* theta, double* Fm, double dt,
                int nfast, hipStream_t st) {
  const LP p = make_lp(s);
  const Div dv = divs(g);
  const int N = g->N, L = s->L;
  const double n_ = N, e_ = g->E, l_ = L;
  if (s->air) {
    ISW_LAUNCH("air_stiff", 3 * l_ * n_ * D, k_air_stiff, c1(N), BS1, N, p, h, theta, s->S);
  } else {
    ISW_LAUNCH("lay_eta", (l_ + 2) * n_ * D, k_lay_eta, c1(N), BS1, N, L, h, s->b, s->eta0);
    ISW_LAUNCH("lay_ubar", l_ * n_ * D + l_ * e_ * D + 2 * e_ * I + (s->wall ? e_ * D : 0) + e_ * D,
               k_lay_ubar, g1(g->E), BS, *g, L, s->shelf, h, u, s->wall, s->ub0);
  }
  if (s->nu4 != 0) {
    lap(g, L, u, s->D1, s->Z1, s->L1, st);
    lap(g, L, s->L1, s->D1, s->Z1, s->L2, st);
  }
  stage(g, s, p, dv, 0, h, u, h, u, 1.0 / 3.0, dt, nfast, s->h1, s->u1, nullptr, theta, st);
  stage(g, s, p, dv, 1, s->h1, s->u1, h, u, 1.0 / 2.0, dt, nfast, s->h2, s->u2, nullptr, theta, st);
  stage(g, s, p, dv, 2, s->h2, s->u2, h, u, 1.0, dt, nfast, h, u, Fm, theta, st);
  if ((s->vmix != 0 || s->vmix_cd != 0) && L > 1) {
    ISW_LAUNCH("lay_vmix", l_ * n_ * D + 2 * e_ * I + 2 * l_ * e_ * D + e_ * D + (s->wall ? e_ * D : 0),
               k_lay_vmix, g1(g->E), BS, *g, L, s->vmix, dt, s->vmix_cd, h, s->wall, u, s->diss);
    if (s->air)
      ISW_LAUNCH("vmix_heat", 6 * n_ * I + 3 * e_ * D + n_ * D + 3 * l_ * n_ * D,
                 k_vmix_heat, c1(N), BS1, *g, p, h, s->diss, theta);
  }
  return (int)hipGetLastError();
}

int layers_lap(const isw_grid* g, const isw_layers* s, const double* u, double* out, hipStream_t st) {
  lap(g, s->L, u, s->D1, s->Z1, out, st);
  return (int)hipGetLastError();
}

int layers_gfast(const isw_grid* g, const isw_layers* s, const double* h, const double* theta, hipStream_t st) {
  if (s->air)
    ISW_LAUNCH("air_gfast", 3.0 * s->L * g->N * D,
               k_air_gfast, c1(g->N), BS1, g->N, make_lp(s), s->rd, h, theta, s->Gc);
  return (int)hipGetLastError();
}

int layers_exchange(const isw_grid* g, const isw_layers* s, const double* wx, double* h, double* u, double* M,
                    double* psi, double* theta, int seq, double nu4dt, hipStream_t st) {
  const int N = g->N, L = s->L;
  const double n_ = N, e_ = g->E, l_ = L;
  ISW_LAUNCH("copy", 2 * l_ * n_ * D, k_copy, g1((long)N * L), BS, (long)N * L, h, s->hold);
  const bool heat = s->air && theta;
  ISW_LAUNCH("xch_cell", (l_ - 1) * n_ * D + n_ * D + 6 * l_ * n_ * D,
             k_xch_cell, c1(N), BS1, N, L, seq, wx, g->areaCell, h, M, psi);
  const bool v4 = heat && s->nu4 != 0;
  ISW_LAUNCH("xch_edge", (l_ - 1) * n_ * D + l_ * n_ * D + 2 * e_ * I + 2 * l_ * e_ * D + (heat ? e_ * D : 0) +
                             (v4 ? l_ * e_ * D : 0) + (s->wall ? e_ * D : 0),
             k_xch_edge, g1(g->E), BS, *g, L, wx, s->hold, u, heat ? s->diss : nullptr,
                                                 v4 ? s->L2 : nullptr, v4 ? nu4dt : 0.0, heat ? s->thin_wind : 0.0,
                                                 s->floor_, s->wall);
  if (heat)
    ISW_LAUNCH("vmix_heat", 6 * n_ * I + 3 * e_ * D + n_ * D + 3 * l_ * n_ * D,
               k_vmix_heat, c1(N), BS1, *g, make_lp(s), h, s->diss, theta);
  return (int)hipGetLastError();
}

int vmix_heat(const isw_grid* g, const isw_layers* s, const double* h, const double* diss, double* theta,
              hipStream_t st) {
  const double n_ = g->N, e_ = g->E, l_ = s->L;
  ISW_LAUNCH("vmix_heat", 6 * n_ * I + 3 * e_ * D + n_ * D + 3 * l_ * n_ * D,
             k_vmix_heat, c1(g->N), BS1, *g, make_lp(s), h, diss, theta);
  return (int)hipGetLastError();
}

int mass_update(const isw_grid* g, int L, const double* M0, const double* Fm, double* M1, hipStream_t st) {
  const double n_ = g->N, e_ = g->E, l_ = L;
  ISW_LAUNCH("mass_update", 2 * l_ * n_ * D + l_ * e_ * D + 6 * n_ * I + e_ * I,
             k_mass_update, c2(g->N, L), BS1, *g, M0, Fm, M1);
  return (int)hipGetLastError();
}

// Transport.step_many (transport.py:40-71) on rows stacked in psi; the result is written back into psi.
int transport_step(const isw_grid* g, const isw_transport* tr, const double* M0, const double* M1, const double* Fm,
                   double* psi, int rows, int order, hipStream_t st) {
  const int N = g->N, E = g->E, Lm = tr->L;
  const double n_ = N, e_ = E, l_ = Lm, r_ = rows, mixed = ISW_FORM_MIXED ? 1 : 0;
#if ISW_FORM_MIXED
  float* pf = tr->pf;
  const fast* wt = g->d2wtf;
#else
  float* pf = nullptr;
  const fast* wt = g->d2wt;
#endif
  const double cellE = 6 * n_ * I + 2 * e_ * I;  // cellEdgeAsc, eI/eJ
  const double sten = 8 * e_ * (R + I) + l_ * e_ * D + 2 * e_ * I;
  ISW_LAUNCH("tr_td", 2 * l_ * n_ * D + r_ * n_ * D + l_ * e_ * D + cellE + r_ * n_ * D + mixed * r_ * n_ * 4,
             k_tr_td, c4(N, rows), dim3(BS1, 4), *g, Lm, rows, M0, psi, Fm, M1, tr->td, pf);
  if (order == 1) {
    ISW_LAUNCH("copy", 2 * r_ * n_ * D, k_copy, g1((long)N * rows), BS, (long)N * rows, tr->td, psi);
    return (int)hipGetLastError();
  }
#if ISW_FORM_MIXED
  const fast* psf = pf;
  const fast* p1f = pf;
  const fast* p2f = pf;
#else
  const fast* psf = psi;
  const fast* p1f = tr->p1;
  const fast* p2f = tr->p2;
#endif
  // (the flux of a patch's edges for four rows in LDS: at most 64 KiB)
  const bool fused = g->pi_off && g->pcells == BS1 && (!ISW_FORM_MIXED || tr->pf2) &&
                     (size_t)4 * g->ptedges * sizeof(double) <= 65536;
  if (fused) {
    // (algorithmic bytes: the stencil's field and its FP32 copy, the stencil and its weights, Fm, M0, psi, the
    // output and its copy, the cell rings and the patch tables; the flux stays in LDS)
    const double tabs = 6 * n_ * I + 2 * (n_ / BS1 + 1) * I + 2 * e_ * I + 6 * n_ * I;
    const size_t shm = (size_t)4 * g->ptedges * sizeof(double);
#if ISW_FORM_MIXED
    float* pfb = tr->pf2;
    const fast* p1s = tr->pf2;                     // (stage 1 writes p1's copy to pf2; stage 2 reads it, writes pf)
#else
    float* pfb = nullptr;
    const fast* p1s = tr->p1;
#endif
    const dim3 gp((unsigned)g->npatch, (unsigned)((rows + 3) / 4));
    ISW_LAUNCH_SHM("tr_hstage", r_ * n_ * D + mixed * r_ * n_ * 4 + sten + l_ * n_ * D + r_ * n_ * D + tabs +
                                    r_ * n_ * D + mixed * r_ * n_ * 4,
                   k_tr_hstage, gp, dim3(BS1, 4), shm, *g, Lm, rows, 1.0 / 3.0, tr->beta, psi, psf, wt, Fm, M0, psi,
                   tr->p1, pfb);
    ISW_LAUNCH_SHM("tr_hstage", r_ * n_ * D + mixed * r_ * n_ * 4 + sten + l_ * n_ * D + r_ * n_ * D + tabs +
                                    r_ * n_ * D + mixed * r_ * n_ * 4,
                   k_tr_hstage, gp, dim3(BS1, 4), shm, *g, Lm, rows, 1.0 / 2.0, tr->beta, tr->p1, p1s, wt, Fm, M0,
                   psi, tr->p2, pf);
  } else {
    ISW_LAUNCH("tr_hflux", r_ * n_ * D + mixed * r_ * n_ * 4 + sten + r_ * e_ * D,
               k_tr_hflux, c4(E, rows), dim3(BS1, 4), *g, Lm, rows, 1.0 / 3.0, tr->beta, psi, psf, wt, Fm, tr->Hf);
    ISW_LAUNCH("tr_stage", l_ * n_ * D + r_ * n_ * D + r_ * e_ * D + l_ * e_ * D + cellE + r_ * n_ * D +
                               mixed * r_ * n_ * 4,
               k_tr_stage, c4(N, rows), dim3(BS1, 4), *g, Lm, rows, 1.0 / 3.0, M0, psi, tr->Hf, Fm, tr->p1, pf);
    ISW_LAUNCH("tr_hflux", r_ * n_ * D + mixed * r_ * n_ * 4 + sten + r_ * e_ * D,
               k_tr_hflux, c4(E, rows), dim3(BS1, 4), *g, Lm, rows, 1.0 / 2.0, tr->beta, tr->p1, p1f, wt, Fm,
               tr->Hf);
    ISW_LAUNCH("tr_stage", l_ * n_ * D + r_ * n_ * D + r_ * e_ * D + l_ * e_ * D + cellE + r_ * n_ * D +
                               mixed * r_ * n_ * 4,
               k_tr_stage, c4(N, rows), dim3(BS1, 4), *g, Lm, rows, 1.0 / 2.0, M0, psi, tr->Hf, Fm, tr->p2, pf);
  }
  ISW_LAUNCH("tr_aflux", 2 * r_ * n_ * D + mixed * r_ * n_ * 4 + sten + r_ * e_ * D,
             k_tr_aflux, dim3((unsigned)((E + BS - 1) / BS), (unsigned)((rows + 3) / 4)), dim3(BS1, 4), *g, Lm, rows,
             tr->beta, psi, tr->p2, p2f, wt, Fm, tr->A);
  ISW_LAUNCH("tr_ratios", 2 * r_ * n_ * D + l_ * n_ * D + r_ * e_ * D + 6 * n_ * I + cellE + 2 * r_ * n_ * D,
             k_tr_ratios, c4(N, rows), dim3(BS1, 4), *g, Lm, rows, psi, tr->td, M1, tr->A, tr->Rp, tr->Rm);
  ISW_LAUNCH("tr_apply", r_ * n_ * D + l_ * n_ * D + r_ * e_ * D + 2 * r_ * n_ * D + cellE + r_ * n_ * D,
             k_tr_apply, c4(N, rows), dim3(BS1, 4), *g, Lm, rows, tr->td, M1, tr->A, tr->Rp, tr->Rm, psi);
  return (int)hipGetLastError();
}

int swe_step(const isw_grid* g, const isw_layers* s, double* h, double* u, double* Fm, double dt, hipStream_t st) {
  const LP p = make_lp(s);
  const Div dv = divs(g);
  const int N = g->N, E = g->E, NV = g->NV, L = s->L;
  const double n_ = N, e_ = E, v_ = NV, l_ = L;
  if (s->nu4 != 0) {
    lap(g, L, u, s->D1, s->Z1, s->L1, st);
    lap(g, L, s->L1, s->D1, s->Z1, s->L2, st);
  }
  const double cs[3] = {dt / 3, dt / 2, dt};
  const double* hsv[3] = {h, s->h1, s->h2};
  const double* usv[3] = {u, s->u1, s->u2};
  double* hnv[3] = {s->h1, s->h2, h};
  double* unv[3] = {s->u1, s->u2, u};
  for (int k = 0; k < 3; k++) {
    const double c = cs[k];
    ISW_LAUNCH("swe_scale", l_ * n_ * D + l_ * e_ * D + 6 * n_ * I + 2 * e_ * I + e_ * D + n_ * D + l_ * n_ * D,
               k_swe_scale, c2(N, L), BS1, *g, c, hsv[k], usv[k], s->scale);
    ISW_LAUNCH("swe_flux", l_ * n_ * D + l_ * e_ * D + 2 * e_ * I + l_ * n_ * D + l_ * e_ * D +
                               (k == 2 && Fm ? l_ * e_ * D + e_ * D : 0),
               k_swe_flux, g2(E, L), BS, *g, k == 2 ? dt : 0.0, hsv[k], usv[k], s->scale, s->F,
                                                   k == 2 ? Fm : nullptr);
    ISW_LAUNCH("swe_cell", 2 * l_ * n_ * D + n_ * D + 2 * l_ * e_ * D + 6 * n_ * (I + C) + 2 * e_ * D + n_ * D +
                               2 * l_ * n_ * D,
               k_swe_cell, c2(N, L), BS1, *g, s->grav, c, hsv[k], s->b, s->F, usv[k], h, hnv[k], s->B);
    ISW_LAUNCH("swe_vertex", l_ * n_ * D + l_ * e_ * D + 3 * v_ * (2 * I + 2 * C) + 6 * n_ * D + e_ * D +
                                 2 * v_ * D + l_ * v_ * D,
               k_swe_vertex, c2(NV, L), BS1, *g, dv, hsv[k], usv[k], s->fV, s->q);
    ISW_LAUNCH("swe_edge", l_ * e_ * D + l_ * v_ * D + l_ * n_ * D + l_ * e_ * D + g->n_eoe * (D + I) +
                               (e_ + 1) * I + 4 * e_ * I + e_ * D + (s->nu4 != 0 ? l_ * e_ * D : 0) + 2 * l_ * e_ * D,
               k_swe_edge_lp, g2(E, L), BS, *g, p, c, s->F, s->q, s->B, usv[k], s->L2, u, unv[k]);
  }
  return (int)hipGetLastError();
}

int couple_stress(const isw_grid* g, int La, const double* ha, double* ua, const double* ho, double* uo, double c,
                  double rho_w, double dt, double* diss, double* tau_acc, const double* wall, double land_drag,
                  const double* cd, hipStream_t st) {
  const double n_ = g->N, e_ = g->E;
  ISW_LAUNCH("couple_stress", 2 * n_ * D + 4 * e_ * D + 2 * e_ * I + e_ * D + 2 * e_ * D + (wall ? e_ * D : 0) +
                                  (cd ? n_ * D : 0),
             k_couple_stress, g1(g->E), BS, *g, La, c, rho_w, dt, land_drag, ha, ua, ho, uo, diss, tau_acc,
                                                      wall, cd);
  return (int)hipGetLastError();
}

int land_column(const isw_grid* g, int K, const double* land, double* T, const double* cap, double* Tg,
                const double* dz, double c_soil, double k_soil, double dt, hipStream_t st) {
  const double n_ = g->N;
  ISW_LAUNCH("land_column", n_ * D + 2 * n_ * D + n_ * D + 2.0 * K * n_ * D + K * D,
             k_land_column, c1(g->N), BS1, g->N, K, c_soil, k_soil, dt, land, T, cap, Tg, dz);
  return (int)hipGetLastError();
}

int ocn_cap(const isw_grid* g, const double* h, double rho_cw, double* cap, const double* land, double c_land,
            hipStream_t st) {
  const double n_ = g->N;
  ISW_LAUNCH("ocn_cap", n_ * D + 2 * n_ * D + (land ? n_ * D : 0),
             k_ocn_cap, c1(g->N), BS1, g->N, rho_cw, c_land, h, cap, land);
  return (int)hipGetLastError();
}

int fill(double* x, long n, double v, hipStream_t st) {
  if (n > 0) ISW_LAUNCH("fill", 8.0 * n, k_fill, g1(n), BS, n, v, x);
  return (int)hipGetLastError();
}

}  // namespace NS
