// The Moon's and the Sun's tides on the ocean (reference: kernels/tides.cu, tides.py:55-87): the tide-generating
// potential of both bodies at every cell, the internal tide's drag per edge from the stratification above the floor,
// and the mixing the bottom stress's dissipation drives. FP64 (the reference's double path), -ffp-contract=off.
//
// Design (MI355X): the bodies' places (two Kepler solves, the Moon's inclined orbit turned to the equator) are the
// same for every cell, so each 256-thread block computes them once on one lane into LDS (the reference recomputes
// them on every thread) and its cells read them from there; the values are the same to the bit whichever lane
// computes them. The potential is then a streaming pass over the cells' positions. The mixing is a column pass (a
// cell's layers on one thread, in the reference's order); the drag an edge pass.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "diag.hpp"
#include "insilica_hip.h"

namespace isw_tides {

constexpr int BS = 256;
constexpr double TPI = 3.141592653589793;

// Kepler's equation E - e sin E = M by six Newton steps from E = M (tides.cu:26-31): the true anomaly, r = 1 - e cos E.
__device__ double kepler(double M, double e, double* r) {
  double E = M;
  for (int it = 0; it < 6; it++) E -= (E - e * sin(E) - M) / (1 - e * cos(E));
  *r = 1 - e * cos(E);
  return 2 * atan2(sqrt(1 + e) * sin(E / 2), sqrt(1 - e) * cos(E / 2));
}

// The Sun's and the Moon's declinations, right ascensions and distances (in semi-major axes) and the day's turn at
// time t (tides.cu:35-56).
struct Sky {
  double decs, ras, rot, rs, decm, ram, rm;
};
__device__ Sky bodies(const isw_sky& tp, double t) {
  Sky s;
  double ls;
  s.rs = 1;
  if (!(tp.ecc > 0)) {
    ls = 2 * TPI * t / tp.year_s;
  } else {
    const double nu0 = -tp.lam_peri, E0 = 2 * atan(sqrt((1 - tp.ecc) / (1 + tp.ecc)) * tan(nu0 / 2));
    ls = kepler(E0 - tp.ecc * sin(E0) + 2 * TPI * t / tp.year_s, tp.ecc, &s.rs) + tp.lam_peri;
  }
  s.ras = atan2(cos(tp.tilt) * sin(ls), cos(ls));
  s.decs = asin(sin(tp.tilt) * sin(ls));
  s.rot = 2 * TPI * fmod(t, tp.day_s) / tp.day_s;
  const double e = tp.moon_e, inc = tp.moon_i;
  const double peri = tp.aps0 + 2 * TPI * t / tp.aps_T, node = tp.node0 + 2 * TPI * t / tp.node_T;
  const double lo = kepler(tp.moon_L0 - peri + 2 * TPI * t / tp.moon_T, e, &s.rm) + peri;
  const double u = lo - node;
  const double bet = asin(sin(inc) * sin(u)), lam = node + atan2(cos(inc) * sin(u), cos(u));
  s.decm = asin(sin(bet) * cos(tp.tilt) + cos(bet) * sin(tp.tilt) * sin(lam));
  s.ram = atan2(sin(lam) * cos(tp.tilt) - tan(bet) * sin(tp.tilt), cos(lam));
  return s;
}

// The potential V = love sum_b GM_b R^2 / r_b^3 (3 cos^2 psi_b - 1) / 2 at each cell (tides.cu:58-71), at time[0] - back.
__global__ __launch_bounds__(BS) void k_tide_pot(int N, isw_sky tp, const double* __restrict__ time, double back,
                                                 const double* __restrict__ pos, double* __restrict__ V) {
  // per body: cos(dec), cos(lon), sin(lon), sin(dec) and GM R^2 / r^3 (the Sun's, then the Moon's)
  __shared__ double dir[10];
  if (threadIdx.x == 0) {
    const Sky s = bodies(tp, time[0] - back);
    const double ls_e = -s.rot, lm_e = -s.rot + (s.ram - s.ras);
    dir[0] = cos(s.decs);
    dir[1] = cos(ls_e);
    dir[2] = sin(ls_e);
    dir[3] = sin(s.decs);
    dir[4] = cos(s.decm);
    dir[5] = cos(lm_e);
    dir[6] = sin(lm_e);
    dir[7] = sin(s.decm);
    const double R2 = tp.radius * tp.radius, as = tp.sun_a * s.rs, am = tp.moon_a * s.rm;
    dir[8] = tp.sun_gm * R2 / (as * as * as);
    dir[9] = tp.moon_gm * R2 / (am * am * am);
  }
  __syncthreads();
  const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (i >= N) return;
  const double x = pos[3 * (size_t)i], y = pos[3 * (size_t)i + 1], z = pos[3 * (size_t)i + 2];
  const double cs = x * dir[0] * dir[1] + y * dir[0] * dir[2] + z * dir[3];
  const double cm = x * dir[4] * dir[5] + y * dir[4] * dir[6] + z * dir[7];
  V[i] = tp.love * (dir[8] * (3 * cs * cs - 1) / 2 + dir[9] * (3 * cm * cm - 1) / 2);
}

// Tidal mixing (tides.cu:85-122): the bottom stress's dissipation D = rho c |u_b|^2 booked in dsum, a share gamma of
// it raising the column's potential energy by mixing heat and salt across each interface, fading as exp(-z / zeta)
// with the interface's height z above the floor; equal volumes exchanged, so heat and salt are kept exactly.
__global__ __launch_bounds__(BS) void k_tide_mix(int N, int E, int L, const double* __restrict__ h,
                                                 const double* __restrict__ u, double* __restrict__ T,
                                                 double* __restrict__ S, const int* __restrict__ cellEdge,
                                                 const double* __restrict__ le, const double* __restrict__ de,
                                                 const double* __restrict__ area, const double* __restrict__ land,
                                                 double c_bot, double rho, double gamma, double dt, isw_ocean o,
                                                 double* __restrict__ dsum, double cd, const double* __restrict__ cbot,
                                                 double zeta) {
  const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (i >= N) return;
  if (land && land[i] > 0.5) return;
  const double* ub = u + (size_t)(L - 1) * E;
  double s = 0, cit = 0;
  int ne = 0;
  for (int m = 0; m < 6; m++) {
    const int e = cellEdge[i * 6 + m];
    if (e < 0) break;
    s += le[e] * de[e] * ub[e] * ub[e];
    if (cbot) cit += cbot[e];
    ne++;
  }
  const double u2 = s / (2 * area[i]);
  const double D = rho * (c_bot + cd * sqrt(u2) + (ne ? cit / ne : 0.0)) * u2;
  dsum[i] += D * dt;
  const double En0 = gamma * D * dt;
  double z = 0;
  for (int k = L - 2; k >= 0; k--) {
    const size_t a = (size_t)k * N + i, b = (size_t)(k + 1) * N + i;
    const double h1 = h[a], h2 = h[b];
    z += h2;
    const double En = En0 * exp(-z / zeta);
    const double t1 = T[a] - 273.15, t2 = T[b] - 273.15;
    const double db = o.grav * (o.alpha * (t1 - t2) + 0.5 * o.alpha2 * (t1 * t1 - t2 * t2) - o.beta * (S[a] - S[b]));
    if (!(db > 0) || !(h1 > 0) || !(h2 > 0)) continue;
    const double full = rho * h1 * h2 * db / 2;
    const double f = En < full ? En / full : 1.0;
    const double v = f * h1 * h2 / (h1 + h2);
    const double dT = T[b] - T[a], dS = S[b] - S[a];
    T[a] += v / h1 * dT;
    T[b] -= v / h2 * dT;
    S[a] += v / h1 * dS;
    S[b] -= v / h2 * dS;
  }
}

// The internal tide's drag per edge (tides.cu:128-144): c = kappa h^2 N / 2 at each end, their mean; N the case's
// (nbot) or from the two lowest layers' buoyancy step over the distance between their middles.
__global__ __launch_bounds__(BS) void k_tide_itd(int N, int E, int L, const double* __restrict__ h,
                                                 const double* __restrict__ T, const double* __restrict__ S,
                                                 const double* __restrict__ h2, const int* __restrict__ eI,
                                                 const int* __restrict__ eJ, double kappa, isw_ocean o,
                                                 double* __restrict__ cbot, const double* __restrict__ nbot) {
  const int e = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (e >= E) return;
  double c = 0;
  for (int s2 = 0; s2 < 2; s2++) {
    const int i = s2 ? eJ[e] : eI[e];
    if (!(h2[i] > 0)) continue;
    if (nbot) {
      c += 0.5 * kappa * h2[i] * nbot[i];
      continue;
    }
    const size_t a = (size_t)(L - 2) * N + i, b = (size_t)(L - 1) * N + i;
    const double t1 = T[a] - 273.15, t2 = T[b] - 273.15;
    const double db = o.grav * (o.alpha * (t1 - t2) + 0.5 * o.alpha2 * (t1 * t1 - t2 * t2) - o.beta * (S[a] - S[b]));
    const double dz = 0.5 * (h[a] + h[b]);
    if (db > 0 && dz > 0) c += 0.5 * kappa * h2[i] * sqrt(db / dz);
  }
  cbot[e] = 0.5 * c;
}

}  // namespace isw_tides

using namespace isw_tides;

namespace {
inline dim3 grid1(int n) { return dim3((unsigned)((n + BS - 1) / BS)); }
bool sky_ok(const isw_sky* tp) {
  return tp && tp->year_s > 0 && tp->day_s > 0 && tp->sun_a > 0 && tp->moon_a > 0 && tp->moon_T != 0 &&
         tp->node_T != 0 && tp->aps_T != 0 && tp->radius > 0 && tp->ecc >= 0 && tp->ecc < 1 && tp->moon_e >= 0 &&
         tp->moon_e < 1;
}
}  // namespace

extern "C" {

int32_t isw_tide_pot(const isw_grid* g, const isw_sky* tp, const double* time, double back, const double* pos,
                     double* V, void* stream) {
  if (!g || g->N <= 0 || !sky_ok(tp) || !time || !pos || !V) return ISW_EBADARG;
  hipStream_t st = (hipStream_t)stream;
  const int N = g->N;
  ISW_LAUNCH("tide_pot", 4.0 * N * 8 + 8, k_tide_pot, grid1(N), dim3(BS), N, *tp, time, back, pos, V);
  return (int32_t)hipGetLastError();
}

int32_t isw_tide_itd(const isw_grid* g, int32_t L, const isw_ocean* o, const double* h, const double* T,
                     const double* S, const double* rough, double kappa, const double* nbot, double* cbot,
                     void* stream) {
  if (!g || g->N <= 0 || g->E <= 0 || !g->eI || !g->eJ || L < 2 || L > ISW_LMAX || !o || !h || !T || !S || !rough ||
      !cbot)
    return ISW_EBADARG;
  hipStream_t st = (hipStream_t)stream;
  const int N = g->N, E = g->E;
  ISW_LAUNCH("tide_itd", (6.0 * N + (nbot ? N : 0) + E) * 8 + 2.0 * E * 4, k_tide_itd, grid1(E), dim3(BS), N, E,
             (int)L, h, T, S, rough, g->eI, g->eJ, kappa, *o, cbot, nbot);
  return (int32_t)hipGetLastError();
}

int32_t isw_tide_mix(const isw_grid* g, int32_t L, const isw_ocean* o, const double* h, const double* u, double* T,
                     double* S, const double* land, double c_bot, double rho, double gamma, double dt, double* dsum,
                     double cd, const double* cbot, double zeta, void* stream) {
  if (!g || g->N <= 0 || g->E <= 0 || !g->cellEdge || !g->le || !g->de || !g->areaCell || L < 2 || L > ISW_LMAX ||
      !o || !h || !u || !T || !S || !dsum || !(rho > 0) || !(zeta > 0) || !(dt > 0))
    return ISW_EBADARG;
  hipStream_t st = (hipStream_t)stream;
  const int N = g->N, E = g->E;
  const double l_ = L;
  ISW_LAUNCH("tide_mix", (l_ * N + 3.0 * E + 4 * l_ * N + (land ? N : 0) + N + 2.0 * N + (cbot ? E : 0)) * 8 +
                             6.0 * N * 4,
             k_tide_mix, grid1(N), dim3(BS), N, E, (int)L, h, u, T, S, g->cellEdge, g->le, g->de, g->areaCell, land,
             c_bot, rho, gamma, dt, *o, dsum, cd, cbot, zeta);
  return (int32_t)hipGetLastError();
}

}  // extern "C"
