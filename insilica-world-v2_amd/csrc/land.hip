// Water over and through the land, and the ice sheets' flow (reference: kernels/land.cu, land.py:15-60), one thread
// per cell on the cell-ordered mesh, FP64.
//
// Surface water (pond, kg/m^2) runs toward lower water surfaces z = b + pond / rho_w + ice / rho_ice through each edge
// at the speed 3/4 U over a depth min(head, the donor's depth), at most 1/12 of the head (times the smaller area) per
// edge; soil water seeps between water tables zt = (soil / soilcap) b by Darcy's law K (le / de) dz, at most 1/12 of
// what levels the two tables; a donor gives at most half its pond and all its soil in a step (the scale pass); a sea
// cell only receives, into mouth. The ice flows by Glen's shallow-ice law (n = 3, no sliding): through an edge
// gam H^5 |grad s|^2 (head / de) le per step, |grad s|^2 the TRiSK kinetic-energy rule's at each end, at most 1/12 of
// the head per edge and all the donor's ice; ice reaching a sea cell becomes its sea ice (calved, ice_prev).
//
// The heights z, zt and s are cell-local, so they are computed where they are read rather than stored: the route takes
// two launches (scale, apply) and the ice three (slope, scale, apply) instead of the reference's three and four, with
// every value as the reference computes it. The new pond, soil and ice go to a work buffer and are copied back.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "diag.hpp"
#include "insilica_hip.h"

namespace isw_land {

constexpr int BS = 256;

inline dim3 grid1(int n) { return dim3((n + BS - 1) / BS); }

struct Route {
  const double *b, *land, *soilcap, *pond, *soil, *ice;
  double rho_w, rho_ice, u, ks;                       // u = uflow dt, ks = kseep dt
};

__device__ __forceinline__ bool is_land(const double* land, int c) { return land[c] > 0.5; }

// the water surface and the water table of cell c (0 at sea)
__device__ __forceinline__ double surf(const Route& r, int c) {
  return is_land(r.land, c) ? r.b[c] + r.pond[c] / r.rho_w + r.ice[c] / r.rho_ice : 0.0;
}
__device__ __forceinline__ double table(const Route& r, int c) {
  return is_land(r.land, c) && r.soilcap[c] > 0 && r.b[c] > 0 ? r.soil[c] / r.soilcap[c] * r.b[c] : 0.0;
}

// The surface flow fp and the seepage fs from cell a to cell c across edge e (kg per step), >= 0 when a gives.
__device__ void flows(const Route& r, const isw_grid& g, int a, int c, int e, double& fp, double& fs) {
  fp = 0;
  fs = 0;
  if (!is_land(r.land, a)) return;
  const bool lc = is_land(r.land, c);
  const double ma = g.areaCell[a] < g.areaCell[c] ? g.areaCell[a] : g.areaCell[c];
  const double za = surf(r, a), head = za - (lc ? surf(r, c) : 0.0);
  if (head > 0 && r.pond[a] > 0) {
    const double d = r.pond[a] / r.rho_w, depth = head < d ? head : d;
    const double q = 0.75 * r.u * depth * g.le[e], cap = head / 12 * ma;
    fp = r.rho_w * (q < cap ? q : cap);
  }
  const double ta = table(r, a), tc = lc ? table(r, c) : 0.0, dz = ta - tc;
  if (dz > 0 && r.soil[a] > 0) {
    const double ra = ta / r.soil[a];                  // the table's height per kg/m^2
    const double rb = lc && r.soilcap[c] > 0 ? (tc > 0 && r.soil[c] > 0 ? tc / r.soil[c] : ra) : 0.0;
    const double q = r.rho_w * r.ks * g.le[e] / g.de[e] * dz, cap = dz / (ra + rb) / 12 * ma;
    fs = q < cap ? q : cap;
  }
}

__device__ __forceinline__ int other(const isw_grid& g, int e, int i) { return g.eI[e] == i ? g.eJ[e] : g.eI[e]; }

// Each land cell's scale: what it would give in all, limited to half its pond and all its soil.
__global__ __launch_bounds__(BS) void k_land_scale(isw_grid g, Route r, double* __restrict__ sp, double* __restrict__ ss) {
  const int i = blockIdx.x * BS + threadIdx.x;
  if (i >= g.N) return;
  sp[i] = 1;
  ss[i] = 1;
  if (!is_land(r.land, i)) return;
  double op = 0, os = 0;
  for (int m = 0; m < 6; m++) {
    const int e = g.cellEdge[i * 6 + m];
    if (e < 0) break;
    double fp, fs;
    flows(r, g, i, other(g, e, i), e, fp, fs);
    op += fp;
    os += fs;
  }
  const double a = g.areaCell[i];
  const double ap = r.pond[i] > 0 ? 0.5 * r.pond[i] * a : 0.0, as = r.soil[i] > 0 ? r.soil[i] * a : 0.0;
  if (op > 0 && op > ap) sp[i] = ap / op;
  if (os > 0 && os > as) ss[i] = as / os;
}

// The new pond and soil from the flows out of this cell and into it; a sea cell adds what it receives to mouth. qout
// (nullable): a land cell's surface outflow, summed (kg/m^2 of its area).
__global__ __launch_bounds__(BS) void k_land_apply(isw_grid g, Route r, const double* __restrict__ sp,
                                                   const double* __restrict__ ss, double* __restrict__ pond1,
                                                   double* __restrict__ soil1, double* __restrict__ mouth,
                                                   double* __restrict__ qout) {
  const int i = blockIdx.x * BS + threadIdx.x;
  if (i >= g.N) return;
  double gp = 0, gs = 0, op = 0;
  const bool li = is_land(r.land, i);
  for (int m = 0; m < 6; m++) {
    const int e = g.cellEdge[i * 6 + m];
    if (e < 0) break;
    const int o = other(g, e, i);
    double fp, fs;
    if (li) {
      flows(r, g, i, o, e, fp, fs);
      gp -= fp * sp[i];
      gs -= fs * ss[i];
      op += fp * sp[i];
    }
    flows(r, g, o, i, e, fp, fs);
    gp += fp * sp[o];
    gs += fs * ss[o];
  }
  const double a = g.areaCell[i];
  if (li) {
    pond1[i] = r.pond[i] + gp / a;
    soil1[i] = r.soil[i] + gs / a;
    if (qout) qout[i] += op / a;
  } else {
    pond1[i] = 0;
    soil1[i] = 0;
    mouth[i] += (gp + gs) / a;
  }
}

struct IceFlow {
  const double *b, *land, *ice;
  double rho_ice, gam;                                 // gam = 2 A (rho_ice g)^3 / 5 dt
};

__device__ __forceinline__ double ice_surf(const IceFlow& f, int c) {
  return is_land(f.land, c) ? f.b[c] + f.ice[c] / f.rho_ice : 0.0;
}

// |grad s|^2 at each cell by the TRiSK kinetic energy's rule: (1 / A) sum_e le de / 2 ((s_o - s_i) / de)^2.
__global__ __launch_bounds__(BS) void k_ice_slope(isw_grid g, IceFlow f, double* __restrict__ g2) {
  const int i = blockIdx.x * BS + threadIdx.x;
  if (i >= g.N) return;
  double sum = 0;
  if (is_land(f.land, i)) {
    const double si = ice_surf(f, i);
    for (int m = 0; m < 6; m++) {
      const int e = g.cellEdge[i * 6 + m];
      if (e < 0) break;
      const double ds = (ice_surf(f, other(g, e, i)) - si) / g.de[e];
      sum += 0.5 * g.le[e] * g.de[e] * ds * ds;
    }
  }
  g2[i] = sum / g.areaCell[i];
}

// The ice from cell a to cell c across edge e (kg per step), >= 0 when a gives; only land gives.
__device__ double ice_flow(const IceFlow& f, const isw_grid& g, const double* g2, int a, int c, int e) {
  if (!is_land(f.land, a) || !(f.ice[a] > 0)) return 0;
  const double head = ice_surf(f, a) - ice_surf(f, c);
  if (!(head > 0)) return 0;
  const double H = f.ice[a] / f.rho_ice, H2 = H * H;
  const double gg = is_land(f.land, c) ? 0.5 * (g2[a] + g2[c]) : g2[a];
  const double q = f.gam * H2 * H2 * H * gg * head / g.de[e] * g.le[e];
  const double ma = g.areaCell[a] < g.areaCell[c] ? g.areaCell[a] : g.areaCell[c], cap = head / 12 * ma;
  return f.rho_ice * (q < cap ? q : cap);
}

__global__ __launch_bounds__(BS) void k_ice_scale(isw_grid g, IceFlow f, const double* __restrict__ g2,
                                                  double* __restrict__ si) {
  const int i = blockIdx.x * BS + threadIdx.x;
  if (i >= g.N) return;
  si[i] = 1;
  if (!is_land(f.land, i)) return;
  double out = 0;
  for (int m = 0; m < 6; m++) {
    const int e = g.cellEdge[i * 6 + m];
    if (e < 0) break;
    out += ice_flow(f, g, g2, i, other(g, e, i), e);
  }
  const double have = f.ice[i] > 0 ? f.ice[i] * g.areaCell[i] : 0.0;
  if (out > have) si[i] = have / out;
}

__global__ __launch_bounds__(BS) void k_ice_apply(isw_grid g, IceFlow f, const double* __restrict__ g2,
                                                  const double* __restrict__ si, double* __restrict__ ice1,
                                                  double* __restrict__ ice_prev, double* __restrict__ calved,
                                                  double* __restrict__ iout) {
  const int i = blockIdx.x * BS + threadIdx.x;
  if (i >= g.N) return;
  double d = 0, out = 0;
  const bool li = is_land(f.land, i);
  for (int m = 0; m < 6; m++) {
    const int e = g.cellEdge[i * 6 + m];
    if (e < 0) break;
    const int o = other(g, e, i);
    if (li) {
      const double fl = ice_flow(f, g, g2, i, o, e) * si[i];
      d -= fl;
      out += fl;
    }
    d += ice_flow(f, g, g2, o, i, e) * si[o];
  }
  d /= g.areaCell[i];
  ice1[i] = f.ice[i] + d;
  if (iout && li) iout[i] += out / g.areaCell[i];
  if (!li) {
    calved[i] += d;
    if (ice_prev) ice_prev[i] += d;
  }
}

__global__ __launch_bounds__(BS) void k_copy2(int n, const double* __restrict__ a, double* __restrict__ x,
                                              const double* __restrict__ b, double* __restrict__ y) {
  const int i = blockIdx.x * BS + threadIdx.x;
  if (i >= n) return;
  x[i] = a[i];
  if (y) y[i] = b[i];
}

int route(const isw_grid* g, const double* b, const double* land, const double* soilcap, double* pond, double* soil,
          const double* ice, double rho_w, double rho_ice, double u_dt, double k_dt, double* work, double* mouth,
          double* qout, hipStream_t st) {
  const int N = g->N;
  const Route r{b, land, soilcap, pond, soil, ice, rho_w, rho_ice, u_dt, k_dt};
  double *sp = work, *ss = work + N, *pond1 = work + 2 * (size_t)N, *soil1 = work + 3 * (size_t)N;
  const double bytes1 = 8.0 * N * (6 + 2 + 2) + 4.0 * N * 6 + 8.0 * g->E * 2 + 8.0 * (g->E * 2);
  ISW_LAUNCH("land_scale", bytes1, k_land_scale, grid1(N), dim3(BS), *g, r, sp, ss);
  ISW_LAUNCH("land_apply", bytes1 + 8.0 * N * 3, k_land_apply, grid1(N), dim3(BS), *g, r, sp, ss, pond1, soil1,
             mouth, qout);
  ISW_LAUNCH("land_copy", 8.0 * N * 4, k_copy2, grid1(N), dim3(BS), N, pond1, pond, soil1, soil);
  return (int)hipGetLastError();
}

int ice(const isw_grid* g, const double* b, const double* land, double* ice_, double rho_ice, double gam_dt,
        double* work, double* ice_prev, double* calved, double* iout, hipStream_t st) {
  const int N = g->N;
  const IceFlow f{b, land, ice_, rho_ice, gam_dt};
  double *g2 = work, *si = work + N, *ice1 = work + 2 * (size_t)N;
  const double bytes1 = 8.0 * N * 4 + 4.0 * N * 6 + 8.0 * g->E * 2 + 8.0 * (g->E * 2);
  ISW_LAUNCH("ice_slope", bytes1, k_ice_slope, grid1(N), dim3(BS), *g, f, g2);
  ISW_LAUNCH("ice_scale", bytes1 + 8.0 * N * 2, k_ice_scale, grid1(N), dim3(BS), *g, f, g2, si);
  ISW_LAUNCH("ice_apply", bytes1 + 8.0 * N * 5, k_ice_apply, grid1(N), dim3(BS), *g, f, g2, si, ice1, ice_prev,
             calved, iout);
  ISW_LAUNCH("ice_copy", 8.0 * N * 2, k_copy2, grid1(N), dim3(BS), N, ice1, ice_, nullptr, nullptr);
  return (int)hipGetLastError();
}

}  // namespace isw_land
