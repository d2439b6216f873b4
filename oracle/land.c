/* The CPU oracle of the land's water routing and ice flow -- TEST INFRASTRUCTURE ONLY.
 *
 * A scalar C restatement of the reference's kernels/land.cu (land_heights, land_scale, land_apply: lines 38-101;
 * ice_surface, ice_slope, ice_scale, ice_apply: 104-173) as LandWater.step and ice_step sequence them (land.py:33-60),
 * in the reference's FP64 arithmetic; the tests check the HIP library against it. */
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

static int landc(const double* land, int c) { return land[c] > 0.5; }
static int other(const og_grid* g, int e, int i) { return g->eI[e] == i ? g->eJ[e] : g->eI[e]; }

/* edge_flows (land.cu:17-35): surface flow fp and seepage fs from a to c across e, from the heights z, zt */
static void edge_flows(const og_grid* g, int a, int c, int e, const double* land, const double* z, const double* pond,
                       const double* zt, const double* soil, const double* soilcap, double rho_w, double u, double ks,
                       double* fp, double* fs) {
  *fp = 0;
  *fs = 0;
  if (!landc(land, a)) return;
  const int lc = landc(land, c);
  const double ma = g->areaCell[a] < g->areaCell[c] ? g->areaCell[a] : g->areaCell[c];
  const double head = z[a] - (lc ? z[c] : 0.0);
  if (head > 0 && pond[a] > 0) {
    const double d = pond[a] / rho_w, depth = head < d ? head : d;
    const double q = 0.75 * u * depth * g->le[e], cap = head / 12 * ma;
    *fp = rho_w * (q < cap ? q : cap);
  }
  const double dz = zt[a] - (lc ? zt[c] : 0.0);
  if (dz > 0 && soil[a] > 0) {
    const double ra = zt[a] / soil[a];
    const double rb = lc && soilcap[c] > 0 ? (zt[c] > 0 && soil[c] > 0 ? zt[c] / soil[c] : ra) : 0.0;
    const double q = rho_w * ks * g->le[e] / g->de[e] * dz, cap = dz / (ra + rb) / 12 * ma;
    *fs = q < cap ? q : cap;
  }
}

void og_land_route(const og_grid* g, const double* b, const double* land, const double* soilcap, double* pond,
                   double* soil, const double* ice, double rho_w, double rho_ice, double u, double ks, double* mouth,
                   double* qout) {
  const int N = g->N;
  double* z = malloc(sizeof(double) * N * 6);
  double *zt = z + N, *sp = z + 2 * N, *ss = z + 3 * N, *pond1 = z + 4 * N, *soil1 = z + 5 * N;
  for (int i = 0; i < N; i++) {                        /* land_heights */
    if (!landc(land, i)) {
      z[i] = zt[i] = 0;
      continue;
    }
    z[i] = b[i] + pond[i] / rho_w + ice[i] / rho_ice;
    zt[i] = soilcap[i] > 0 && b[i] > 0 ? soil[i] / soilcap[i] * b[i] : 0.0;
  }
  for (int i = 0; i < N; i++) {                        /* land_scale */
    sp[i] = ss[i] = 1;
    if (!landc(land, i)) continue;
    double op = 0, os = 0;
    for (int m = 0; m < 6; m++) {
      const int e = g->cellEdge[i * 6 + m];
      if (e < 0) break;
      double fp, fs;
      edge_flows(g, i, other(g, e, i), e, land, z, pond, zt, soil, soilcap, rho_w, u, ks, &fp, &fs);
      op += fp;
      os += fs;
    }
    const double ap = pond[i] > 0 ? 0.5 * pond[i] * g->areaCell[i] : 0.0;
    const double as = soil[i] > 0 ? soil[i] * g->areaCell[i] : 0.0;
    if (op > 0 && op > ap) sp[i] = ap / op;
    if (os > 0 && os > as) ss[i] = as / os;
  }
  for (int i = 0; i < N; i++) {                        /* land_apply */
    double gp = 0, gs = 0, op = 0;
    const int li = landc(land, i);
    for (int m = 0; m < 6; m++) {
      const int e = g->cellEdge[i * 6 + m];
      if (e < 0) break;
      const int o = other(g, e, i);
      double fp, fs;
      if (li) {
        edge_flows(g, i, o, e, land, z, pond, zt, soil, soilcap, rho_w, u, ks, &fp, &fs);
        gp -= fp * sp[i];
        gs -= fs * ss[i];
        op += fp * sp[i];
      }
      edge_flows(g, o, i, e, land, z, pond, zt, soil, soilcap, rho_w, u, ks, &fp, &fs);
      gp += fp * sp[o];
      gs += fs * ss[o];
    }
    if (li) {
      pond1[i] = pond[i] + gp / g->areaCell[i];
      soil1[i] = soil[i] + gs / g->areaCell[i];
      if (qout) qout[i] += op / g->areaCell[i];
    } else {
      pond1[i] = soil1[i] = 0;
      mouth[i] += (gp + gs) / g->areaCell[i];
    }
  }
  memcpy(pond, pond1, sizeof(double) * N);
  memcpy(soil, soil1, sizeof(double) * N);
  free(z);
}

/* ice_flow (land.cu:136-150) */
static double ice_flow(const og_grid* g, int a, int c, int e, const double* land, const double* s, const double* g2,
                       const double* ice, double rho_ice, double gam) {
  if (!landc(land, a) || !(ice[a] > 0)) return 0;
  const double head = s[a] - s[c];
  if (!(head > 0)) return 0;
  const double H = ice[a] / rho_ice, H2 = H * H;
  const double gg = landc(land, c) ? 0.5 * (g2[a] + g2[c]) : g2[a];
  const double q = gam * H2 * H2 * H * gg * head / g->de[e] * g->le[e];
  const double ma = g->areaCell[a] < g->areaCell[c] ? g->areaCell[a] : g->areaCell[c], cap = head / 12 * ma;
  return rho_ice * (q < cap ? q : cap);
}

void og_ice_flow(const og_grid* g, const double* b, const double* land, double* ice, double rho_ice, double gam,
                 double* ice_prev, double* calved, double* iout) {
  const int N = g->N;
  double* s = malloc(sizeof(double) * N * 4);
  double *g2 = s + N, *si = s + 2 * N, *ice1 = s + 3 * N;
  for (int i = 0; i < N; i++) s[i] = landc(land, i) ? b[i] + ice[i] / rho_ice : 0.0;
  for (int i = 0; i < N; i++) {
    double sum = 0;
    if (landc(land, i))
      for (int m = 0; m < 6; m++) {
        const int e = g->cellEdge[i * 6 + m];
        if (e < 0) break;
        const double ds = (s[other(g, e, i)] - s[i]) / g->de[e];
        sum += 0.5 * g->le[e] * g->de[e] * ds * ds;
      }
    g2[i] = sum / g->areaCell[i];
  }
  for (int i = 0; i < N; i++) {
    si[i] = 1;
    if (!landc(land, i)) continue;
    double out = 0;
    for (int m = 0; m < 6; m++) {
      const int e = g->cellEdge[i * 6 + m];
      if (e < 0) break;
      out += ice_flow(g, i, other(g, e, i), e, land, s, g2, ice, rho_ice, gam);
    }
    const double have = ice[i] > 0 ? ice[i] * g->areaCell[i] : 0.0;
    if (out > have) si[i] = have / out;
  }
  for (int i = 0; i < N; i++) {
    double d = 0, out = 0;
    const int li = landc(land, i);
    for (int m = 0; m < 6; m++) {
      const int e = g->cellEdge[i * 6 + m];
      if (e < 0) break;
      const int o = other(g, e, i);
      if (li) {
        const double f = ice_flow(g, i, o, e, land, s, g2, ice, rho_ice, gam) * si[i];
        d -= f;
        out += f;
      }
      d += ice_flow(g, o, i, e, land, s, g2, ice, rho_ice, gam) * si[o];
    }
    d /= g->areaCell[i];
    ice1[i] = ice[i] + d;
    if (iout && li) iout[i] += out / g->areaCell[i];
    if (!li) {
      calved[i] += d;
      if (ice_prev) ice_prev[i] += d;
    }
  }
  memcpy(ice, ice1, sizeof(double) * N);
  free(s);
}
