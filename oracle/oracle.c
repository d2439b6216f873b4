/* CPU oracle -- TEST INFRASTRUCTURE ONLY (parity checker and CPU baseline; never linked into the product).
 *
 * A scalar restatement of the reference's double-precision kernels, each function citing the reference kernel it
 * follows (file:line under src/insilica/), with the same operation order, so that the HIP library's plain FP64 form
 * can be compared with it value for value. Layered arrays are layer-major x[l * n + i].
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define FOR_N(i, n) for (int i = 0; i < (n); i++)
#define PAR _Pragma("omp parallel for schedule(static)")

struct og_layers {
  og_grid g;
  og_layers_cfg c;
  double dmin;
  double *Gc, *B, *q, *Tp, *F, *scale, *G, *He, *ubs, *ub0, *ub, *acc, *eta0, *eta, *eta_prev, *h1, *h2, *u1, *u2;
  double *D1, *Z1, *L1, *L2, *hold, *hedge, *diss, *pibar, *S, *muf, *mup, *phi, *uf, *accL, *pk;
  double nu4dt;
};

static double* zalloc(size_t n) { return (double*)calloc(n ? n : 1, sizeof(double)); }

/* ---- mesh operators (kernels/trisk.cu:14-45) ---- */

static void cell_div(const og_grid* g, int L, const double* u, double* D) {
  for (int l = 0; l < L; l++) {
    const double* ul = u + (size_t)l * g->E;
    PAR FOR_N(i, g->N) {
      double s = 0;
      for (int k = 0; k < 6; k++) {
        const int e = g->cellEdge[i * 6 + k];
        if (e < 0) break;
        s += (double)g->cellSign[i * 6 + k] * g->le[e] * ul[e];
      }
      D[(size_t)l * g->N + i] = s / g->areaCell[i];
    }
  }
}

static void vertex_curl(const og_grid* g, int L, const double* u, double* Z) {
  for (int l = 0; l < L; l++) {
    const double* ul = u + (size_t)l * g->E;
    PAR FOR_N(v, g->NV) {
      double z = 0;
      for (int k = 0; k < 3; k++) {
        const int e = g->vEdge[v * 3 + k];
        z += (double)g->vEdgeSign[v * 3 + k] * g->de[e] * ul[e];
      }
      Z[(size_t)l * g->NV + v] = z / g->areaV[v];
    }
  }
}

static void edge_lap(const og_grid* g, int L, const double* D, const double* Z, double* out) {
  for (int l = 0; l < L; l++) {
    const double* Dl = D + (size_t)l * g->N;
    const double* Zl = Z + (size_t)l * g->NV;
    PAR FOR_N(e, g->E) {
      out[(size_t)l * g->E + e] = (Dl[g->eJ[e]] - Dl[g->eI[e]]) / g->de[e] - (Zl[g->ev2[e]] - Zl[g->ev1[e]]) / g->le[e];
    }
  }
}

/* swe_vertex (kernels/fused.cu:82-98): q = (zeta + f) / h_v, kite-weighted h_v, in the 64-bit form (DIVBY divides). */
static void swe_vertex(const og_grid* g, int L, const double* h, const double* u, const double* fV, double* q) {
  for (int l = 0; l < L; l++) {
    const double* hl = h + (size_t)l * g->N;
    const double* ul = u + (size_t)l * g->E;
    PAR FOR_N(v, g->NV) {
      double hv = 0, z = 0;
      for (int k = 0; k < 3; k++) {
        const int c = g->vCell[v * 3 + k];
        hv += g->kite[c * 6 + g->vSlot[v * 3 + k]] * hl[c] * 1.0;
      }
      for (int k = 0; k < 3; k++) {
        const int e = g->vEdge[v * 3 + k];
        z += (double)g->vEdgeSign[v * 3 + k] * g->de[e] * ul[e];
      }
      q[(size_t)l * g->NV + v] = (z / g->areaV[v] + fV[v]) / (hv / g->areaV[v]);
    }
  }
}

/* ---- the layered core (kernels/layers.cu) ---- */

/* edge_h (layers.cu:42-45) with ob == NULL (shelf_at gives shelf). */
static inline double edge_h(double a, double b, double shelf) {
  const double m = a < b ? a : b;
  return shelf > 0 && m < shelf ? m : 0.5 * (a + b);
}

static void lay_mask(const og_grid* g, const double* wall, int L, double* a) {
  if (!wall) return;
  PAR FOR_N(e, g->E) {
    if (wall[e] == 0)
      for (int k = 0; k < L; k++) a[(size_t)k * g->E + e] = 0;
  }
}

/* lay_eta (layers.cu:33-38) */
static void lay_eta(const og_grid* g, int L, const double* h, const double* b, double* eta) {
  PAR FOR_N(i, g->N) {
    double z = b[i];
    for (int k = L - 1; k >= 0; k--) z += h[(size_t)k * g->N + i];
    eta[i] = z;
  }
}

/* lay_ubar (layers.cu:64-74) */
static void lay_ubar(const og_grid* g, int L, const double* h, const double* u, double shelf, double* ubar) {
  const int N = g->N;
  PAR FOR_N(e, g->E) {
    double H = 0, U = 0;
    for (int k = 0; k < L; k++) {
      const double he = edge_h(h[(size_t)k * N + g->eI[e]], h[(size_t)k * N + g->eJ[e]], shelf);
      H += he;
      U += he * u[(size_t)k * g->E + e];
    }
    ubar[e] = U / H;
  }
}

static inline double cell_ke(const og_grid* g, const double* ul, int i) {
  double ke = 0;
  for (int m = 0; m < 6; m++) {
    const int e = g->cellEdge[i * 6 + m];
    if (e < 0) break;
    ke += g->le[e] * g->de[e] * 0.25 * ul[e] * ul[e];
  }
  return ke;
}

/* lay_bern (layers.cu:83-109), sig NULL; tide (N) the tide-generating potential or NULL */
static void lay_bern(const og_grid* g, int L, const double* h, const double* u, const double* b, const double* gp,
                     const double* tide, double* B) {
  const int N = g->N;
  PAR FOR_N(i, N) {
    double z[OG_LMAX];
    double s = b[i];
    for (int k = L - 1; k >= 0; k--) { s += h[(size_t)k * N + i]; z[k] = s; }
    double Mp = 0;
    for (int k = 0; k < L; k++) {
      if (k > 0) Mp += gp[k - 1] * z[k];
      const double ke = cell_ke(g, u + (size_t)k * g->E, i);
      double Bk = Mp + ke / g->areaCell[i];
      if (tide) Bk -= tide[i];
      B[(size_t)k * N + i] = Bk;
    }
  }
}

/* air_bern (layers.cu:118-158) */
static void air_bern(const og_grid* g, int L, const double* h, const double* u, const double* theta, const double* b,
                     double grav, double cp, double kappa, double p0, const double* S, double* B, double* pibar) {
  const int N = g->N;
  PAR FOR_N(i, N) {
    double xb[OG_LMAX];
    double p = 0, mu = 0, xt = 0;
    for (int k = 0; k < L; k++) {
      const double hk = h[(size_t)k * N + i];
      const double pt = p;
      p = p + grav * hk;
      mu += hk;
      const double x = pow(p / p0, kappa);
      xb[k] = x;
      pibar[(size_t)k * N + i] = cp * (p * x - pt * xt) / ((kappa + 1) * (p - pt));
      xt = x;
    }
    (void)mu;
    double M = grav * b[i] + theta[(size_t)(L - 1) * N + i] * cp * xb[L - 1];
    double Mk[OG_LMAX];
    Mk[L - 1] = M;
    for (int k = L - 2; k >= 0; k--) {
      M = M + (theta[(size_t)k * N + i] - theta[(size_t)(k + 1) * N + i]) * cp * xb[k];
      Mk[k] = M;
    }
    double P[OG_LMAX], Phi[OG_LMAX];
    double a = 0;
    for (int k = 0; k < L; k++) { a += h[(size_t)k * N + i]; P[k] = a; }
    double ph = 0;
    for (int k = L - 1; k >= 0; k--) { ph += S[(size_t)k * N + i] * P[k]; Phi[k] = ph; }
    for (int k = 0; k < L; k++) {
      const double ke = cell_ke(g, u + (size_t)k * g->E, i);
      B[(size_t)k * N + i] = Mk[k] - Phi[k] + ke / g->areaCell[i];
    }
  }
}

/* air_stiff (layers.cu:165-175) */
static void air_stiff(const og_grid* g, int L, const double* h, const double* theta, double grav, double cp,
                      double kappa, double p0, double* S) {
  const int N = g->N;
  PAR FOR_N(i, N) {
    double p = 0;
    for (int k = 0; k < L; k++) {
      p += grav * h[(size_t)k * N + i];
      const double dth = theta[(size_t)k * N + i] - (k + 1 < L ? theta[(size_t)(k + 1) * N + i] : 0.0);
      const double dpi = kappa * cp * pow(p / p0, kappa) / p;
      S[(size_t)k * N + i] = dth > 0 ? dth * grav * dpi : 0.0;
    }
  }
}

/* air_gfast (layers.cu:234-250) */
static void air_gfast(const og_grid* g, int L, const double* h, const double* theta, double grav, double kappa,
                      double p0, double rgas, double floor_, double* Gc) {
  const int N = g->N;
  PAR FOR_N(i, N) {
    double mu = 0;
    for (int k = 0; k < L; k++) mu += h[(size_t)k * N + i];
    int kb = L - 1;
    while (kb > 0 && h[(size_t)kb * N + i] < 2 * floor_) kb--;
    const double Ts = theta[(size_t)kb * N + i] * pow(grav * mu / p0, kappa);
    double gc = rgas * Ts / mu;
    double above = 0;
    for (int j = 0; j < L - 1; j++) {
      above += h[(size_t)j * N + i];
      const double dth = theta[(size_t)j * N + i] - theta[(size_t)(j + 1) * N + i];
      if (dth > 0) gc += dth * rgas * pow(grav * above / p0, kappa) * above / (mu * mu);
    }
    Gc[i] = gc;
  }
}

/* af_muphi (layers.cu:182-207) */
static void af_muphi(const og_grid* g, int L, const double* hedge, const double* uf, double tau, double* mu,
                     double* mu_prev, double gamma, const double* S, double* phi) {
  const int N = g->N, E = g->E;
  PAR FOR_N(i, N) {
    double P[OG_LMAX];
    double a = 0;
    for (int k = 0; k < L; k++) {
      const double* hl = hedge + (size_t)k * E;
      const double* ul = uf + (size_t)k * E;
      const size_t o = (size_t)k * N + i;
      const double m0 = mu[o];
      mu_prev[o] = m0;
      double div = 0;
      for (int m = 0; m < 6; m++) {
        const int e = g->cellEdge[i * 6 + m];
        if (e < 0) break;
        div += (double)g->cellSign[i * 6 + m] * g->le[e] * hl[e] * ul[e];
      }
      const double m1 = m0 - tau * div / g->areaCell[i];
      mu[o] = m1;
      a += m1 + gamma * (m1 - m0);
      P[k] = a;
    }
    double ph = 0;
    for (int k = L - 1; k >= 0; k--) { ph += S[(size_t)k * N + i] * P[k]; phi[(size_t)k * N + i] = ph; }
  }
}

/* af_u (layers.cu:209-217) */
static void af_u(const og_grid* g, int L, const double* phi, const double* Tp, const double* G, double tau, double* uf,
                 double* acc) {
  const int N = g->N, E = g->E;
  for (int l = 0; l < L; l++) {
    PAR FOR_N(e, E) {
      const size_t o = (size_t)l * E + e;
      const double v = uf[o];
      acc[o] += v;
      const double* pl = phi + (size_t)l * N;
      uf[o] = v + tau * (-(pl[g->eJ[e]] - pl[g->eI[e]]) / g->de[e] + Tp[o] + G[e]);
    }
  }
}

/* af_flux (layers.cu:220-224) */
static void af_flux(const og_grid* g, int L, const double* hedge, const double* acc, double inv_n, double* F) {
  const size_t n = (size_t)L * g->E;
  PAR for (size_t o = 0; o < n; o++) F[o] = hedge[o] * acc[o] * inv_n;
}

/* lay_edge_hq (layers.cu:254-263) */
static void lay_edge_hq(const og_grid* g, int L, const double* h, const double* q, double shelf, const double* u,
                        double* he, double* pk) {
  const int N = g->N, E = g->E, NV = g->NV;
  for (int l = 0; l < L; l++) {
    PAR FOR_N(e, E) {
      const size_t o = (size_t)l * E + e;
      const double hv = edge_h(h[(size_t)l * N + g->eI[e]], h[(size_t)l * N + g->eJ[e]], shelf);
      const double qv = 0.5 * (q[(size_t)l * NV + g->ev1[e]] + q[(size_t)l * NV + g->ev2[e]]);
      he[o] = hv;
      double* pp = pk + ((size_t)e * L + l) * 3;
      pp[0] = hv; pp[1] = u[o]; pp[2] = qv;
    }
  }
}

/* lay_slow (layers.cu:269-359) with sig, zm, ob NULL; the bottom layer's quadratic drag cd_bot and the internal
 * tide's drag cbot (E or NULL) as the reference's tides set them */
static void lay_slow(const og_grid* g, int L, const double* pk, const double* hedge, const double* u, const double* B,
                     const double* drag, const double* stress, double floor_, double nu4, const double* lap4,
                     const double* theta, const double* pibar, double dts, double thin, double shelf, double* Tp,
                     double* G, double* He, double* ubs, double cd_bot, const double* cbot) {
  const int N = g->N, E = g->E;
  PAR FOR_N(e, E) {
    double T[OG_LMAX], hev[OG_LMAX], sqk[OG_LMAX];
    double H = 0, U = 0, GT = 0;
    const int eI = g->eI[e], eJ = g->eJ[e];
    for (int k = 0; k < L; k++) {
      const double* hl = hedge + (size_t)k * E;
      const double* ul = u + (size_t)k * E;
      const double* Bl = B + (size_t)k * N;
      const double he = hl[e];
      const double cap = thin * he, qf = pk[((size_t)e * L + k) * 3 + 2];
      double Q = 0;
      for (int m = g->eoeStart[e]; m < g->eoeStart[e + 1]; m++) {
        const double* pp = pk + ((size_t)g->eoe[m] * L + k) * 3;
        double h2 = pp[0];
        if (h2 > cap) h2 = cap;
        const double F2 = h2 * pp[1];
        Q += g->w[m] * F2 * 0.5 * (qf + pp[2]);
      }
      const double shk = shelf;
      const double sq0 = shk > 0 && he < shk ? (he - 2 * floor_ > 0 ? (he - 2 * floor_) / (shk - 2 * floor_) : 0.0) : 1.0;
      const double sq = theta ? 1.0 : sq0;
      sqk[k] = sq0;
      double d = Q - sq * (Bl[eJ] - Bl[eI]) / g->de[e] - (drag[k] + (1 - sq) / 21600.0) * ul[e];
      double ck = stress[k];
      if (k == L - 1 && (cd_bot > 0 || cbot)) {
        /* |u| from the edge's normal wind and its TRiSK tangential wind (the packed neighbours' u) */
        if (cd_bot > 0) {
          double vt = 0;
          for (int m = g->eoeStart[e]; m < g->eoeStart[e + 1]; m++) vt += g->w[m] * pk[((size_t)g->eoe[m] * L + k) * 3 + 1];
          ck += cd_bot * sqrt(ul[e] * ul[e] + vt * vt);
        }
        if (cbot) ck += cbot[e];
      }
      if (ck != 0) {
        const double r = ck / (he > floor_ ? he : floor_);
        d = d - ((r * dts > 1e-8) ? (1 - exp(-r * dts)) / dts : r) * ul[e];
      }
      if (theta) {
        const double* tl = theta + (size_t)k * N;
        const double* pl = pibar + (size_t)k * N;
        d = d + 0.5 * (pl[eI] + pl[eJ]) * (tl[eJ] - tl[eI]) / g->de[e];
      }
      if (nu4 != 0) d = d - nu4 * lap4[(size_t)k * E + e];
      T[k] = d;
      hev[k] = he;
      H += he;
      U += he * ul[e];
    }
    if (theta && shelf > 0) {
      double Tn[OG_LMAX];
      for (int k = 0; k < L; k++) {
        Tn[k] = T[k];
        if (sqk[k] >= 1) continue;
        double w = 0, s = 0;
        for (int j = k - 1; j >= 0; j--) if (sqk[j] >= 1) { w += hev[j]; s += hev[j] * T[j]; break; }
        for (int j = k + 1; j < L; j++) if (sqk[j] >= 1) { w += hev[j]; s += hev[j] * T[j]; break; }
        if (w > 0) Tn[k] = sqk[k] * T[k] + (1 - sqk[k]) * s / w;
      }
      for (int k = 0; k < L; k++) T[k] = Tn[k];
    }
    for (int k = 0; k < L; k++) GT += hev[k] * T[k];
    const double gg = GT / H;
    for (int k = 0; k < L; k++) Tp[(size_t)k * E + e] = T[k] - gg;
    G[e] = gg;
    He[e] = H;
    ubs[e] = U / H;
  }
}

/* bt_eta (layers.cu:366-377) */
static void bt_eta(const og_grid* g, const double* He, const double* ub, double tau, double* eta, double* eta_prev) {
  PAR FOR_N(i, g->N) {
    eta_prev[i] = eta[i];
    double div = 0;
    for (int m = 0; m < 6; m++) {
      const int e = g->cellEdge[i * 6 + m];
      if (e < 0) break;
      div += (double)g->cellSign[i * 6 + m] * g->le[e] * He[e] * ub[e];
    }
    eta[i] = eta[i] - tau * div / g->areaCell[i];
  }
}

/* bt_u (layers.cu:379-388) */
static void bt_u(const og_grid* g, const double* eta, const double* eta_prev, double gamma, const double* G,
                 const double* Gc, double tau, double* ub, double* acc) {
  PAR FOR_N(e, g->E) {
    const double v = ub[e];
    acc[e] += v;
    const int i = g->eI[e], j = g->eJ[e];
    const double ej = Gc[j] * (eta[j] + gamma * (eta[j] - eta_prev[j]));
    const double ei = Gc[i] * (eta[i] + gamma * (eta[i] - eta_prev[i]));
    ub[e] = v + tau * (-(ej - ei) / g->de[e] + G[e]);
  }
}

/* lay_flux (layers.cu:392-397) */
static void lay_flux(const og_grid* g, int L, const double* h, const double* u, const double* ubs, const double* acc,
                     double inv_n, double shelf, double* F) {
  const int N = g->N, E = g->E;
  for (int l = 0; l < L; l++) {
    PAR FOR_N(e, E) {
      const double he = edge_h(h[(size_t)l * N + g->eI[e]], h[(size_t)l * N + g->eJ[e]], shelf);
      F[(size_t)l * E + e] = he * (u[(size_t)l * E + e] - ubs[e]) + he * (acc[e] * inv_n);
    }
  }
}

/* lay_scale (layers.cu:401-416) */
static void lay_scale(const og_grid* g, int L, const double* h0, const double* F, double dts, double floor_,
                      double* scale) {
  const int N = g->N, E = g->E;
  for (int l = 0; l < L; l++) {
    const double* Fl = F + (size_t)l * E;
    PAR FOR_N(i, N) {
      double out = 0;
      for (int m = 0; m < 6; m++) {
        const int e = g->cellEdgeAsc[i * 6 + m];
        if (e < 0) break;
        const double f = Fl[e] * g->le[e] * dts;
        if (g->eI[e] == i) { if (f > 0) out += f; }
        else if (!(f > 0)) out -= f;
      }
      const double avail = h0[(size_t)l * N + i] - floor_;
      const double cap = (avail > 0 ? avail : 0.0) * g->areaCell[i];
      scale[(size_t)l * N + i] = out > cap ? cap / out : 1.0;
    }
  }
}

/* lay_limit (layers.cu:418-426) */
static void lay_limit(const og_grid* g, int L, const double* scale, double dt, double* F, double* Fm) {
  const int N = g->N, E = g->E;
  for (int l = 0; l < L; l++) {
    PAR FOR_N(e, E) {
      const size_t o = (size_t)l * E + e;
      double f = F[o];
      f *= f > 0 ? scale[(size_t)l * N + g->eI[e]] : scale[(size_t)l * N + g->eJ[e]];
      F[o] = f;
      if (Fm) Fm[o] = f * g->le[e] * dt;
    }
  }
}

/* lay_h (layers.cu:429-441) */
static void lay_h(const og_grid* g, int L, const double* h0, const double* F, double cdt, double* hn) {
  const int N = g->N, E = g->E;
  for (int l = 0; l < L; l++) {
    const double* Fl = F + (size_t)l * E;
    PAR FOR_N(i, N) {
      double div = 0;
      for (int m = 0; m < 6; m++) {
        const int e = g->cellEdge[i * 6 + m];
        if (e < 0) break;
        div += (double)g->cellSign[i * 6 + m] * g->le[e] * Fl[e];
      }
      const size_t o = (size_t)l * N + i;
      hn[o] = h0[o] - cdt * div / g->areaCell[i];
    }
  }
}

/* lay_u (layers.cu:443-447) */
static void lay_u(const og_grid* g, int L, const double* u0, const double* Tp, const double* ub_end, const double* ub0,
                  double cdt, double* un) {
  const int E = g->E;
  for (int l = 0; l < L; l++) {
    PAR FOR_N(e, E) {
      const size_t o = (size_t)l * E + e;
      un[o] = u0[o] + cdt * Tp[o] + (ub_end[e] - ub0[e]);
    }
  }
}

/* lay_vmix (layers.cu:575-593) */
static void lay_vmix(const og_grid* g, int L, const double* h, double c, double dt, double* u, double* diss, double cd) {
  const int N = g->N, E = g->E;
  PAR FOR_N(e, E) {
    double d = 0;
    for (int k = 0; k + 1 < L; k++) {
      const double ha = 0.5 * (h[(size_t)k * N + g->eI[e]] + h[(size_t)k * N + g->eJ[e]]);
      const double hb = 0.5 * (h[(size_t)(k + 1) * N + g->eI[e]] + h[(size_t)(k + 1) * N + g->eJ[e]]);
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
  }
}

/* lay_vmix on its own, for the tests' pins of the interfacial stress */
void og_lay_vmix(const og_grid* g, int L, const double* h, double c, double dt, double* u, double* diss, double cd) {
  lay_vmix(g, L, h, c, dt, u, diss, cd);
}

/* vmix_heat (layers.cu:599-621) */
void og_vmix_heat(const og_grid* g, int L, const double* h, const double* diss, double grav, double cp, double kappa,
                  double p0, double* theta) {
  const int N = g->N;
  PAR FOR_N(i, N) {
    double q = 0;
    int skip = 0;
    for (int m = 0; m < 6; m++) {
      const int e = g->cellEdge[i * 6 + m];
      if (e < 0) break;
      q += 0.5 * g->le[e] * g->de[e] * diss[e];
    }
    if (q == 0) skip = 1;
    if (!skip) {
      q /= g->areaCell[i];
      double mu = 0;
      for (int k = 0; k < L; k++) mu += h[(size_t)k * N + i];
      double p = 0, xt = 0;
      for (int k = 0; k < L; k++) {
        const double pt = p;
        p = p + grav * h[(size_t)k * N + i];
        const double x = pow(p / p0, kappa);
        const double pib = cp * (p * x - pt * xt) / ((kappa + 1) * (p - pt));
        theta[(size_t)k * N + i] += q / (mu * pib);
        xt = x;
      }
    }
  }
}

/* xch_cell (layers.cu:454-492) */
static void xch_cell(const og_grid* g, int L, const double* wx, double* h, double* M, double* psi, int seq) {
  const int N = g->N;
  const double* areaCell = g->areaCell;
  PAR FOR_N(i, N) {
    double tr[OG_LMAX], dm[OG_LMAX];
    for (int k = 0; k < L; k++) { tr[k] = 0; dm[k] = 0; }
    if (seq) {
      double m[OG_LMAX], t[OG_LMAX];
      for (int k = 0; k < L; k++) { const size_t o = (size_t)k * N + i; m[k] = M[o]; t[k] = M[o] * psi[o]; }
      for (int pass = 0; pass < 2; pass++)
        for (int n = 0; n + 1 < L; n++) {
          const int k = pass == 0 ? L - 2 - n : n;
          const double v = wx[(size_t)k * N + i];
          if (pass == 0 ? !(v > 0) : !(v < 0)) continue;
          const int d = v > 0 ? k + 1 : k;
          const double vol = v * areaCell[i], p = t[d] / m[d];
          m[k] += vol; m[k + 1] -= vol; t[k] += vol * p; t[k + 1] -= vol * p;
        }
      for (int k = 0; k < L; k++) {
        const size_t o = (size_t)k * N + i;
        h[o] = h[o] + (m[k] - M[o]) / areaCell[i];
        M[o] = m[k];
        psi[o] = t[k] / m[k];
      }
      continue;
    }
    for (int k = 0; k + 1 < L; k++) {
      const double v = wx[(size_t)k * N + i];
      const double p = v > 0 ? psi[(size_t)(k + 1) * N + i] : psi[(size_t)k * N + i];
      const double vol = v * areaCell[i];
      dm[k] += vol; dm[k + 1] -= vol;
      tr[k] += vol * p; tr[k + 1] -= vol * p;
    }
    for (int k = 0; k < L; k++) {
      const size_t o = (size_t)k * N + i;
      const double m1 = M[o] + dm[k];
      psi[o] = (M[o] * psi[o] + tr[k]) / m1;
      M[o] = m1;
      h[o] = h[o] + dm[k] / areaCell[i];
    }
  }
}

/* xch_edge (layers.cu:511-565) */
static void xch_edge(const og_grid* g, int L, const double* wx, const double* hold, double* u, double* diss,
                     const double* lap4, double nu4dt, double hthin, double floor_) {
  const int N = g->N, E = g->E;
  PAR FOR_N(e, E) {
    double mom[OG_LMAX], hn[OG_LMAX];
    double k0 = 0, v4 = 0;
    const int eI = g->eI[e], eJ = g->eJ[e];
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
    double un[OG_LMAX];
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
    for (int k = 0; k < L; k++) {
      u[(size_t)k * E + e] = un[k];
      k1 += hn[k] * un[k] * un[k];
    }
    if (diss) diss[e] = 0.5 * (k0 - k1) + nu4dt * v4;
  }
}

/* ---- Layers (src/insilica/layers.py) ---- */

void og_layers_tides(og_layers* s, const double* tide, const double* cbot, double cd_bot) {
  s->c.tide = tide;
  s->c.cbot = cbot;
  s->c.cd_bot = cd_bot;
}

og_layers* og_layers_new(const og_grid* g, const og_layers_cfg* cfg) {
  og_layers* s = (og_layers*)calloc(1, sizeof(og_layers));
  s->g = *g;
  s->c = *cfg;
  const size_t N = g->N, E = g->E, NV = g->NV, L = cfg->L;
  s->Gc = zalloc(N);
  for (size_t i = 0; i < N; i++) s->Gc[i] = cfg->grav;
  double* const* dummy = 0;
  (void)dummy;
  s->B = zalloc(L * N); s->q = zalloc(L * NV); s->Tp = zalloc(L * E); s->F = zalloc(L * E); s->scale = zalloc(L * N);
  s->G = zalloc(E); s->He = zalloc(E); s->ubs = zalloc(E); s->ub0 = zalloc(E); s->ub = zalloc(E); s->acc = zalloc(E);
  s->eta0 = zalloc(N); s->eta = zalloc(N); s->eta_prev = zalloc(N);
  s->h1 = zalloc(L * N); s->h2 = zalloc(L * N); s->u1 = zalloc(L * E); s->u2 = zalloc(L * E);
  s->D1 = zalloc(L * N); s->Z1 = zalloc(L * NV); s->L1 = zalloc(L * E); s->L2 = zalloc(L * E);
  s->hold = zalloc(L * N); s->hedge = zalloc(L * E); s->diss = zalloc(E);
  s->pibar = zalloc(L * N); s->S = zalloc(L * N); s->muf = zalloc(L * N); s->mup = zalloc(L * N); s->phi = zalloc(L * N);
  s->uf = zalloc(L * E); s->accL = zalloc(L * E); s->pk = zalloc(3 * L * E);
  double dmin = g->de[0];
  for (int e = 1; e < g->E; e++) if (g->de[e] < dmin) dmin = g->de[e];
  s->dmin = dmin;
  return s;
}

void og_layers_free(og_layers* s) {
  if (!s) return;
  double* arrs[] = {s->Gc, s->B, s->q, s->Tp, s->F, s->scale, s->G, s->He, s->ubs, s->ub0, s->ub, s->acc, s->eta0,
                    s->eta, s->eta_prev, s->h1, s->h2, s->u1, s->u2, s->D1, s->Z1, s->L1, s->L2, s->hold, s->hedge,
                    s->diss, s->pibar, s->S, s->muf, s->mup, s->phi, s->uf, s->accL, s->pk};
  for (size_t k = 0; k < sizeof(arrs) / sizeof(arrs[0]); k++) free(arrs[k]);
  free(s);
}

double* og_layers_Gc(og_layers* s) { return s->Gc; }
double* og_layers_L2(og_layers* s) { return s->L2; }
double* og_layers_diss(og_layers* s) { return s->diss; }

/* Layers.update_gfast (layers.py:230-235) */
void og_layers_gfast(og_layers* s, const double* h, const double* theta) {
  if (s->c.air) air_gfast(&s->g, s->c.L, h, theta, s->c.grav, s->c.kappa, s->c.p0, s->c.rd, s->c.floor_, s->Gc);
}

/* Layers.substeps (layers.py:225-228): ceil(dt sqrt(max_i sum_k h_k Gc) / (0.5 d_min)), at least 1. The column sum
 * runs over the layers in order, as cupy's sum over axis 0 of a (L, N) array does for small L. */
int og_layers_substeps(og_layers* s, const double* h, double dt) {
  const int N = s->g.N, L = s->c.L;
  double c2 = -INFINITY;
  for (int i = 0; i < N; i++) {
    double col = 0;
    for (int k = 0; k < L; k++) col += h[(size_t)k * N + i];
    const double v = col * s->Gc[i];
    if (v > c2) c2 = v;
  }
  const double x = ceil(dt * sqrt(c2) / (0.5 * s->dmin));
  return x < 1 ? 1 : (int)x;
}

void og_layers_lap(og_layers* s, const double* u, double* out) {
  cell_div(&s->g, s->c.L, u, s->D1);
  vertex_curl(&s->g, s->c.L, u, s->Z1);
  edge_lap(&s->g, s->c.L, s->D1, s->Z1, out);
}

/* Layers._stage (layers.py:255-309) */
static void stage(og_layers* s, const double* hs, const double* us, const double* h0, const double* u0, double c,
                  double dt, int nfast, double* hn, double* un, double* Fm, const double* theta) {
  const og_grid* g = &s->g;
  const og_layers_cfg* C = &s->c;
  const int L = C->L;
  const size_t LN = (size_t)L * g->N, LE = (size_t)L * g->E;
  if (C->air)
    air_bern(g, L, hs, us, theta, C->b, C->grav, C->cp, C->kappa, C->p0, s->S, s->B, s->pibar);
  else
    lay_bern(g, L, hs, us, C->b, C->gp, C->tide, s->B);
  swe_vertex(g, L, hs, us, C->fV, s->q);
  lay_edge_hq(g, L, hs, s->q, C->shelf, us, s->hedge, s->pk);
  lay_slow(g, L, s->pk, s->hedge, us, s->B, C->drag, C->stress, C->floor_, C->nu4, C->nu4 != 0 ? s->L2 : 0,
           C->air ? theta : 0, C->air ? s->pibar : 0, c * dt, C->thin, C->shelf, s->Tp, s->G, s->He, s->ubs,
           C->air ? 0.0 : C->cd_bot, C->air ? 0 : C->cbot);
  int n = (int)ceil(c * nfast);
  if (n < 1) n = 1;
  const double tau = c * dt / n;
  if (C->air) {
    memcpy(s->muf, h0, LN * sizeof(double));
    memcpy(s->uf, u0, LE * sizeof(double));
    memset(s->accL, 0, LE * sizeof(double));
    for (int it = 0; it < n; it++) {
      af_muphi(g, L, s->hedge, s->uf, tau, s->muf, s->mup, C->div_damp, s->S, s->phi);
      af_u(g, L, s->phi, s->Tp, s->G, tau, s->uf, s->accL);
    }
    af_flux(g, L, s->hedge, s->accL, 1.0 / n, s->F);
    lay_scale(g, L, h0, s->F, c * dt, C->floor_, s->scale);
    lay_limit(g, L, s->scale, dt, s->F, Fm);
    lay_h(g, L, h0, s->F, c * dt, hn);
    memcpy(un, s->uf, LE * sizeof(double));
    return;
  }
  memcpy(s->eta, s->eta0, g->N * sizeof(double));
  memcpy(s->ub, s->ub0, g->E * sizeof(double));
  memset(s->acc, 0, g->E * sizeof(double));
  for (int it = 0; it < n; it++) {
    bt_eta(g, s->He, s->ub, tau, s->eta, s->eta_prev);
    bt_u(g, s->eta, s->eta_prev, C->div_damp, s->G, s->Gc, tau, s->ub, s->acc);
    lay_mask(g, C->wall, 1, s->ub);
    lay_mask(g, C->wall, 1, s->acc);
  }
  lay_flux(g, L, hs, us, s->ubs, s->acc, 1.0 / n, C->shelf, s->F);
  lay_mask(g, C->wall, L, s->F);
  lay_scale(g, L, h0, s->F, c * dt, C->floor_, s->scale);
  lay_limit(g, L, s->scale, dt, s->F, Fm);
  lay_h(g, L, h0, s->F, c * dt, hn);
  lay_u(g, L, u0, s->Tp, s->ub, s->ub0, c * dt, un);
  lay_mask(g, C->wall, L, un);
}

/* Layers.step (layers.py:311-338) */
void og_layers_step(og_layers* s, double* h, double* u, double dt, int nfast, double* Fm, double* theta) {
  const og_grid* g = &s->g;
  const og_layers_cfg* C = &s->c;
  const int L = C->L;
  if (C->air) {
    air_stiff(g, L, h, theta, C->grav, C->cp, C->kappa, C->p0, s->S);
  } else {
    lay_eta(g, L, h, C->b, s->eta0);
    lay_ubar(g, L, h, u, C->shelf, s->ub0);
    lay_mask(g, C->wall, 1, s->ub0);
  }
  if (C->nu4 != 0) {
    og_layers_lap(s, u, s->L1);
    og_layers_lap(s, s->L1, s->L2);
    s->nu4dt = C->nu4 * dt;
  }
  stage(s, h, u, h, u, 1.0 / 3.0, dt, nfast, s->h1, s->u1, 0, theta);
  stage(s, s->h1, s->u1, h, u, 1.0 / 2.0, dt, nfast, s->h2, s->u2, 0, theta);
  stage(s, s->h2, s->u2, h, u, 1.0, dt, nfast, h, u, Fm, theta);
  if ((C->vmix != 0 || C->vmix_cd != 0) && L > 1) {
    lay_vmix(g, L, h, C->vmix, dt, u, s->diss, C->vmix_cd);
    lay_mask(g, C->wall, L, u);
    if (C->air) og_vmix_heat(g, L, h, s->diss, C->grav, C->cp, C->kappa, C->p0, theta);
  }
}

/* Layers.exchange (layers.py:340-358) */
void og_layers_exchange(og_layers* s, const double* wx, double* h, double* u, double* M, double* psi, double* theta,
                        int seq, double nu4dt) {
  const og_grid* g = &s->g;
  const og_layers_cfg* C = &s->c;
  const int L = C->L;
  memcpy(s->hold, h, (size_t)L * g->N * sizeof(double));
  const int heat = C->air && theta;
  xch_cell(g, L, wx, h, M, psi, seq);
  const int v4 = heat && C->nu4 != 0;
  xch_edge(g, L, wx, s->hold, u, heat ? s->diss : 0, v4 ? s->L2 : 0, v4 ? nu4dt : 0.0, heat ? C->thin_wind : 0.0,
           C->floor_);
  lay_mask(g, C->wall, L, u);
  if (heat) og_vmix_heat(g, L, h, s->diss, C->grav, C->cp, C->kappa, C->p0, theta);
}

/* ---- transport (kernels/trisk.cu:48-59 mass_update; kernels/fused.cu:124-241; transport.py:40-71) ---- */

void og_mass_update(const og_grid* g, int L, const double* M0, const double* Fm, double* M1) {
  const int N = g->N, E = g->E;
  for (int l = 0; l < L; l++) {
    const double* f = Fm + (size_t)l * E;
    PAR FOR_N(i, N) {
      const size_t o = (size_t)l * N + i;
      double m = M0[o];
      for (int k = 0; k < 6; k++) {
        const int e = g->cellEdgeAsc[i * 6 + k];
        if (e < 0) break;
        if (g->eI[e] == i) m -= f[e]; else m += f[e];
      }
      M1[o] = m;
    }
  }
}

static inline double upwind(const double* Fml, const double* p, const og_grid* g, int e) {
  const double f = Fml[e];
  return f * (f >= 0 ? p[g->eI[e]] : p[g->eJ[e]]);
}

static inline double hoflux(const double* Fml, const double* ps, const og_grid* g, int e, double beta) {
  const double f = Fml[e];
  const int o = e * 8;
  const int* idx = g->d2idx;
  const double* wt = g->d2wt;
  const double a = wt[o] * ps[idx[o]] + wt[o + 1] * ps[idx[o + 1]] + wt[o + 2] * ps[idx[o + 2]] + wt[o + 3] * ps[idx[o + 3]];
  const double b = wt[o + 4] * ps[idx[o + 4]] + wt[o + 5] * ps[idx[o + 5]] + wt[o + 6] * ps[idx[o + 6]] + wt[o + 7] * ps[idx[o + 7]];
  return f * (0.5 * (ps[g->eI[e]] + ps[g->eJ[e]]) - (a + b)) + fabs(f) * ((b - a) * beta);
}

void og_transport(const og_grid* g, int L, int rows, double beta, const double* M0, const double* M1,
                  const double* Fm, double* psi, int order) {
  const int N = g->N, E = g->E;
  const size_t RN = (size_t)rows * N, RE = (size_t)rows * E;
  double *td = zalloc(RN), *p1 = zalloc(RN), *p2 = zalloc(RN), *Rp = zalloc(RN), *Rm = zalloc(RN), *Hf = zalloc(RE),
         *A = zalloc(RE);
  /* tr_td (fused.cu:146-160) */
  for (int y = 0; y < rows; y++) {
    const int l = y % L;
    const double* p = psi + (size_t)y * N;
    const double* Fml = Fm + (size_t)l * E;
    PAR FOR_N(i, N) {
      const size_t o = (size_t)y * N + i, om = (size_t)l * N + i;
      double t = M0[om] * psi[o];
      for (int k = 0; k < 6; k++) {
        const int e = g->cellEdgeAsc[i * 6 + k];
        if (e < 0) break;
        const double f = upwind(Fml, p, g, e);
        if (g->eI[e] == i) t -= f; else t += f;
      }
      td[o] = t / M1[om];
    }
  }
  if (order == 1) {
    memcpy(psi, td, RN * sizeof(double));
  } else {
    const double fr[2] = {1.0 / 3.0, 1.0 / 2.0};
    const double* src[2] = {psi, p1};
    double* dst[2] = {p1, p2};
    for (int s = 0; s < 2; s++) {
      const double frac = fr[s];
      for (int y = 0; y < rows; y++) { /* tr_hflux (fused.cu:163-167) */
        const int l = y % L;
        PAR FOR_N(e, E) {
          Hf[(size_t)y * E + e] = frac * hoflux(Fm + (size_t)l * E, src[s] + (size_t)y * N, g, e, beta);
        }
      }
      for (int y = 0; y < rows; y++) { /* tr_stage (fused.cu:170-186) */
        const int l = y % L;
        const double* Hl = Hf + (size_t)y * E;
        const double* Fml = Fm + (size_t)l * E;
        PAR FOR_N(i, N) {
          const size_t o = (size_t)y * N + i, om = (size_t)l * N + i;
          double t = M0[om] * psi[o];
          double mx = M0[om];
          for (int k = 0; k < 6; k++) {
            const int e = g->cellEdgeAsc[i * 6 + k];
            if (e < 0) break;
            const double f = Hl[e];
            const double m = frac * Fml[e];
            if (g->eI[e] == i) { t -= f; mx -= m; } else { t += f; mx += m; }
          }
          dst[s][o] = t / mx;
        }
      }
    }
    for (int y = 0; y < rows; y++) { /* tr_aflux (fused.cu:189-194) */
      const int l = y % L;
      const double* Fml = Fm + (size_t)l * E;
      PAR FOR_N(e, E) {
        A[(size_t)y * E + e] = hoflux(Fml, p2 + (size_t)y * N, g, e, beta) - upwind(Fml, psi + (size_t)y * N, g, e);
      }
    }
    for (int y = 0; y < rows; y++) { /* tr_ratios (fused.cu:197-222) */
      const int l = y % L;
      const double* p = psi + (size_t)y * N;
      const double* t = td + (size_t)y * N;
      const double* Al = A + (size_t)y * E;
      PAR FOR_N(i, N) {
        double hi = fmax(p[i], t[i]), lo = fmin(p[i], t[i]);
        for (int k = 0; k < 6; k++) {
          const int j = g->nbr[i * 6 + k];
          if (j < 0) break;
          hi = fmax(hi, fmax(p[j], t[j]));
          lo = fmin(lo, fmin(p[j], t[j]));
        }
        const double m1 = M1[(size_t)l * N + i];
        const double qmax = (hi - t[i]) * m1, qmin = (t[i] - lo) * m1;
        double Pp = 0, Pm = 0;
        for (int k = 0; k < 6; k++) {
          const int e = g->cellEdgeAsc[i * 6 + k];
          if (e < 0) break;
          const double a = Al[e];
          if (g->eI[e] == i) { if (a > 0) Pm += a; else Pp -= a; }
          else { if (a > 0) Pp += a; else Pm -= a; }
        }
        Rp[(size_t)y * N + i] = Pp > 0 ? fmin(1.0, qmax / Pp) : 0.0;
        Rm[(size_t)y * N + i] = Pm > 0 ? fmin(1.0, qmin / Pm) : 0.0;
      }
    }
    for (int y = 0; y < rows; y++) { /* tr_apply (fused.cu:225-241), into psi (the reference's out, then copied) */
      const int l = y % L;
      const double* Al = A + (size_t)y * E;
      const double* rp = Rp + (size_t)y * N;
      const double* rm = Rm + (size_t)y * N;
      PAR FOR_N(i, N) {
        const size_t o = (size_t)y * N + i, om = (size_t)l * N + i;
        double t = td[o] * M1[om];
        for (int k = 0; k < 6; k++) {
          const int e = g->cellEdgeAsc[i * 6 + k];
          if (e < 0) break;
          const double a = Al[e];
          const double c = a > 0 ? fmin(rp[g->eJ[e]], rm[g->eI[e]]) : fmin(rp[g->eI[e]], rm[g->eJ[e]]);
          if (g->eI[e] == i) t -= c * a; else t += c * a;
        }
        p1[o] = t / M1[om];
      }
    }
    memcpy(psi, p1, RN * sizeof(double));
  }
  free(td); free(p1); free(p2); free(Rp); free(Rm); free(Hf); free(A);
}

/* ---- SWE (kernels/fused.cu:27-120; swe.py:105-132) ---- */

void og_swe_step(const og_grid* g, int L, double grav, const double* b, const double* drag, double nu4, double* h,
                 double* u, double dt, double* Fm) {
  const int N = g->N, E = g->E, NV = g->NV;
  const size_t LN = (size_t)L * N, LE = (size_t)L * E;
  double *F = zalloc(LE), *B = zalloc(LN), *q = zalloc((size_t)L * NV), *scale = zalloc(LN), *h1 = zalloc(LN),
         *h2 = zalloc(LN), *u1 = zalloc(LE), *u2 = zalloc(LE), *D1 = zalloc(LN), *Z1 = zalloc((size_t)L * NV),
         *L1 = zalloc(LE), *L2 = zalloc(LE);
  if (nu4 != 0) {
    cell_div(g, L, u, D1); vertex_curl(g, L, u, Z1); edge_lap(g, L, D1, Z1, L1);
    cell_div(g, L, L1, D1); vertex_curl(g, L, L1, Z1); edge_lap(g, L, D1, Z1, L2);
  }
  const double cs[3] = {dt / 3, dt / 2, dt};
  const double* hsv[3] = {h, h1, h2};
  const double* usv[3] = {u, u1, u2};
  double* hnv[3] = {h1, h2, h};
  double* unv[3] = {u1, u2, u};
  for (int s = 0; s < 3; s++) {
    const double c = cs[s];
    const double *hs = hsv[s], *us = usv[s];
    double *hn = hnv[s], *un = unv[s];
    double* Fms = s == 2 ? Fm : 0;
    const double dtm = s == 2 ? dt : 0.0;
    for (int l = 0; l < L; l++) { /* swe_scale (fused.cu:32-48) */
      const double* hl = hs + (size_t)l * N;
      const double* ul = us + (size_t)l * E;
      PAR FOR_N(i, N) {
        double out = 0;
        for (int k = 0; k < 6; k++) {
          const int e = g->cellEdgeAsc[i * 6 + k];
          if (e < 0) break;
          const double f = 0.5 * (hl[g->eI[e]] + hl[g->eJ[e]]) * ul[e] * g->le[e] * c;
          if (g->eI[e] == i) { if (f > 0) out += f; }
          else if (!(f > 0)) out -= f;
        }
        const double hv = hl[i];
        const double cap = 0.5 * (hv > 0 ? hv : 0.0) * g->areaCell[i];
        scale[(size_t)l * N + i] = out > cap ? cap / out : 1.0;
      }
    }
    for (int l = 0; l < L; l++) { /* swe_flux (fused.cu:51-59) */
      const double* hl = hs + (size_t)l * N;
      const double* ul = us + (size_t)l * E;
      PAR FOR_N(e, E) {
        const size_t o = (size_t)l * E + e;
        double f = 0.5 * (hl[g->eI[e]] + hl[g->eJ[e]]) * ul[e];
        f *= f > 0 ? scale[(size_t)l * N + g->eI[e]] : scale[(size_t)l * N + g->eJ[e]];
        F[o] = f;
        if (Fms) Fms[o] = f * g->le[e] * dtm;
      }
    }
    for (int l = 0; l < L; l++) { /* swe_cell (fused.cu:62-79) */
      const double* Fl = F + (size_t)l * E;
      const double* ul = us + (size_t)l * E;
      PAR FOR_N(i, N) {
        double div = 0, ke = 0;
        for (int k = 0; k < 6; k++) {
          const int e = g->cellEdge[i * 6 + k];
          if (e < 0) break;
          div += (double)g->cellSign[i * 6 + k] * g->le[e] * Fl[e];
          ke += g->le[e] * g->de[e] * 0.25 * ul[e] * ul[e];
        }
        const size_t o = (size_t)l * N + i;
        const double dh = 0.0 - div / g->areaCell[i];
        B[o] = grav * (hs[o] + b[i]) + 0.0 + ke / g->areaCell[i];
        hn[o] = h[o] + c * dh;
      }
    }
    swe_vertex(g, L, hs, us, g->fV, q);
    for (int l = 0; l < L; l++) { /* swe_edge (fused.cu:101-120) */
      const double* Fl = F + (size_t)l * E;
      const double* ql = q + (size_t)l * NV;
      const double* Bl = B + (size_t)l * N;
      PAR FOR_N(e, E) {
        const double qa = 0.5 * (ql[g->ev1[e]] + ql[g->ev2[e]]);
        double Q = 0;
        for (int k = g->eoeStart[e]; k < g->eoeStart[e + 1]; k++) {
          const int e2 = g->eoe[k];
          const double qb = 0.5 * (ql[g->ev1[e2]] + ql[g->ev2[e2]]);
          Q += g->w[k] * Fl[e2] * 0.5 * (qa + qb);
        }
        const size_t o = (size_t)l * E + e;
        double d = Q - (Bl[g->eJ[e]] - Bl[g->eI[e]]) / g->de[e] - drag[l] * us[o];
        if (nu4 != 0) d = d - nu4 * L2[o];
        un[o] = u[o] + c * d;
      }
    }
  }
  free(F); free(B); free(q); free(scale); free(h1); free(h2); free(u1); free(u2); free(D1); free(Z1); free(L1); free(L2);
}

/* ---- coupling and ground (kernels/ocean.cu) ---- */

/* couple_stress (ocean.cu:17-42) */
void og_couple_stress(const og_grid* g, int La, const double* ha, double* ua, const double* ho, double* uo, double c0,
                      double rho_w, double dt, double* diss, double* tau_acc, const double* wall, double land_drag0,
                      const double* cd) {
  const int N = g->N, E = g->E;
  PAR FOR_N(e, E) {
    double c = c0, land_drag = land_drag0;
    const int b = La - 1;
    if (cd) { c = 0.5 * (cd[g->eI[e]] + cd[g->eJ[e]]); land_drag = 1; }
    const double ma = 0.5 * (ha[(size_t)b * N + g->eI[e]] + ha[(size_t)b * N + g->eJ[e]]);
    double* pa = ua + (size_t)b * E + e;
    double* po = uo + e;
    const double a0 = *pa;
    if (wall && wall[e] == 0) {
      const double du1 = a0 * exp(-land_drag * c / ma * dt);
      *pa = du1;
      diss[e] = 0.5 * ma * (a0 * a0 - du1 * du1);
    } else {
      const double mo = rho_w * 0.5 * (ho[g->eI[e]] + ho[g->eJ[e]]);
      const double mean = (ma * *pa + mo * *po) / (ma + mo);
      const double du0 = *pa - *po;
      const double du1 = du0 * exp(-c * (1 / ma + 1 / mo) * dt);
      *pa = mean + mo / (ma + mo) * du1;
      *po = mean - ma / (ma + mo) * du1;
      diss[e] = 0.5 * ma * mo / (ma + mo) * (du0 * du0 - du1 * du1);
    }
    const double tau = ma * (a0 - *pa) / dt;
    tau_acc[e] += tau * tau * dt;
  }
}

/* land_column (ocean.cu:434-465) */
void og_land_column(const og_grid* g, int K, const double* land, double* T, const double* cap, double* Tg,
                    const double* dz, double c_soil, double k_soil, double dt) {
  const int N = g->N;
  PAR FOR_N(i, N) {
    if (!(land[i] > 0.5)) continue;
    const int n = K + 1;
    double C[17], G[17], x[17], cp_[17], dp[17];
    C[0] = cap[i]; x[0] = T[i];
    double half = cap[i] / c_soil / 2;
    for (int k = 1; k < n; k++) {
      const double d = dz[k - 1];
      C[k] = c_soil * d; x[k] = Tg[(size_t)(k - 1) * N + i];
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
}

/* ocn_cap (ocean.cu:468-471) */
void og_ocn_cap(const og_grid* g, const double* h, double rho_cw, double* cap, const double* land, double c_land) {
  PAR FOR_N(i, g->N) {
    cap[i] = land && land[i] > 0.5 ? (cap[i] > 0 ? cap[i] : c_land) : rho_cw * h[i];
  }
}

/* block_sum passes (trisk.cu:66-81; gpu.py:111-126): 256 threads x 8 values per block, halving tree. */
double og_dsum(const double* x, const double* y, int64_t n) {
  double* cur = 0;
  const double* xx = x;
  const double* yy = y;
  for (;;) {
    const int64_t blocks = (n + 256 * 8 - 1) / 256 / 8;
    double* out = zalloc((size_t)blocks);
    for (int64_t bk = 0; bk < blocks; bk++) {
      double s[256];
      const int64_t base = bk * 256 * 8;
      for (int t = 0; t < 256; t++) {
        double acc = 0;
        for (int k = 0; k < 8; k++) {
          const int64_t i = base + t + (int64_t)k * 256;
          if (i < n) acc += yy ? xx[i] * yy[i] : xx[i];
        }
        s[t] = acc;
      }
      for (int h = 128; h > 0; h >>= 1)
        for (int t = 0; t < h; t++) s[t] += s[t + h];
      out[bk] = s[0];
    }
    free(cur);
    cur = out;
    if (blocks == 1) {
      const double r = out[0];
      free(out);
      return r;
    }
    xx = out; yy = 0; n = blocks;
  }
}
