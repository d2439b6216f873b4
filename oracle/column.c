/* CPU oracle of the air column's Held-Suarez path and the surface-exchange kernels -- TEST INFRASTRUCTURE ONLY
 * (parity checker and CPU baseline; never linked into the product).
 *
 * A scalar restatement of the reference's double path (src/insilica/kernels/air.cu): hs_column 2065-2109 with the
 * interface crossing it calls (crossing 744-850, move_mass 495-520, column_energy 484-488, edge_density 525-539,
 * edge_shape 245-270 (the double build's), layer_exner 129-140, p_of 186), surface_layer 2015-2061, cell_speed
 * 1978-1989, cell_speed_perot 1994-2006, gas_src 2133-2137, advance_time 1975; in the reference's operation order,
 * built with gcc -O2 -ffp-contract=off.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

enum { P_GRAV = 13, P_CP = 14, P_KAPPA = 15, P_P0 = 16, P_FLOOR = 11, P_THETA_STAR = 10, P_LV = 23, P_MOIST = 22,
       P_MU_BL = 9, P_BL_MIX = 70, P_BL_CAP = 73,
       P_RHO_ICE = 29, P_ICE_COVER = 30, P_BULK_GUST = 75, P_Q_SEA = 97, P_Z0_LAND = 91, P_THS = 132,
       P_EDGE_VAP = 86, P_FREEZE_DRY = 85, P_EDGE_K = 101, P_ICE_AIR = 102, P_SAT_PROF = 51, P_SAT_BASE = 53,
       P_BL_Q = 89, P_Q_SHAPE = 83, P_PLUME_ROOT = 88, P_ROOT_Q = 107, P_PLUME_DEEP = 47, P_PLUME_ENT = 78,
       P_ENT_CLOUD = 93, P_ENT_DET = 108, P_L_DET = 65, P_STORM_AREA = 80, P_GRAUPEL_V = 81, P_ICE_SMALL = 82,
       P_ML_CLEAR = 118, P_CU_CLOUD = 119, P_ENT_IFS = 109, P_ENT_LOC = 125, P_DET_ALL = 116, P_BUOY_V = 113,
       P_SHALLOW = 110, P_CU_DTH = 111, P_CU_DP = 112, P_CU_SUB = 114, P_CL_FLOAT = 117, P_SUB_MIX = 131 };

static void layer_exner(int L, const double* mu, const double* par, double* pb, double* pib) {
  const double grav = par[P_GRAV], p0 = par[P_P0], kappa = par[P_KAPPA], cp = par[P_CP];
  double p = 0, xt = 0;
  for (int k = 0; k < L; k++) {
    const double pt = p;
    p = p + grav * mu[k];
    pb[k] = p;
    const double x = pow(p / p0, kappa);
    pib[k] = cp * (p * x - pt * xt) / ((kappa + 1) * (p - pt));
    xt = x;
  }
}

/* theta* of interface j (between layers j and j+1): the lowest is theta_star, the others par's list from the top. */
static double tstar(int j, int L, const double* par) { return j == L - 2 ? par[P_THETA_STAR] : par[P_THS + j]; }

/* The truncated-exponential shape s of a layer's theta between its edges a < b with mean m: r(s) = (m - a) / (b - a),
 * r(s) = 1 - 1/s + 1/(e^s - 1), by safeguarded Newton in a bracket (air.cu:245-270). */
static double edge_shape(double a, double b, double m) {
  const double W = b - a;
  double x = (m - a) / W;
  if (x < 1e-6) x = 1e-6;
  if (x > 1 - 1e-6) x = 1 - 1e-6;
  double lo = -700, hi = 700, s = 12 * (x - 0.5);
  for (int it = 0; it < 40; it++) {
    double r, dr;
    if (fabs(s) < 1e-4) {
      r = 0.5 + s / 12;
      dr = 1.0 / 12;
    } else {
      const double e1 = expm1(s);
      r = 1 - 1 / s + 1 / e1;
      dr = 1 / (s * s) - (e1 + 1) / (e1 * e1);
    }
    if (r < x) lo = s; else hi = s;
    double sn = s - (r - x) / dr;
    if (r == x) break;
    if (!(sn >= lo && sn <= hi) || !(dr > 0)) sn = 0.5 * (lo + hi);
    if (fabs(sn - s) < 1e-12 * (1 + fabs(s))) {
      s = sn;
      break;
    }
    s = sn;
  }
  return s;
}

/* The layer's share of mass per kelvin at its lower (fa) and upper (fb) edge (air.cu:525-539). */
static void edge_density(double a, double b, double m, double* fa, double* fb) {
  if (!(b > a)) {
    *fa = 0;
    *fb = 0;
    return;
  }
  const double W = b - a, s = edge_shape(a, b, m);
  if (fabs(s) < 1e-6) {
    *fa = (1 - s / 2) / W;
    *fb = (1 + s / 2) / W;
    return;
  }
  const double e1 = expm1(s);
  *fa = s / e1 / W;
  *fb = s * (e1 + 1) / e1 / W;
}

static double column_energy(int L, const double* mu, const double* pib, const double* th, const double* q,
                            const double* par) {
  double E = 0;
  for (int k = 0; k < L; k++) E += mu[k] * pib[k] * th[k] + par[P_LV] * mu[k] * q[k];
  return E;
}

/* Move m (> 0) kg/m^2 from layer `from` to `to`, re-dividing the column; the energy residual to the receiver. */
/* The column's gas tracers (OG_GASN x OG_LMAX mole fractions, xfer) while a column is stepped, else NULL: every move of
 * its air carries them, the donor's fraction into the receiver's (air.cu:503-507). One per OpenMP thread. */
static _Thread_local double* xf_col;

static void move_mass(int L, const double* par, double* mu, double* th, double* q, double* pib, int from, int to,
                      double m, double th_leave, double th_arrive, double q_leave, double q_arrive, double* cl) {
  if (m == 0) return;
  double pb2[OG_LMAX], pib2[OG_LMAX], mu2[OG_LMAX], th2[OG_LMAX], q2[OG_LMAX];
  const double Eb = column_energy(L, mu, pib, th, q, par);
  for (int k = 0; k < L; k++) {
    mu2[k] = mu[k];
    th2[k] = th[k];
    q2[k] = q[k];
  }
  mu2[from] -= m;
  mu2[to] += m;
  if (xf_col) {
    const int d = m > 0 ? from : to, r = m > 0 ? to : from;
    const double am = fabs(m);
    for (int g = 0; g < OG_GASN; g++)
      xf_col[g * OG_LMAX + r] = (mu[r] * xf_col[g * OG_LMAX + r] + am * xf_col[g * OG_LMAX + d]) / mu2[r];
  }
  layer_exner(L, mu2, par, pb2, pib2);
  th2[from] = (mu[from] * th[from] - m * th_leave) / mu2[from];
  th2[to] = (mu[to] * th[to] + m * th_arrive) / mu2[to];
  q2[from] = (mu[from] * q[from] - m * q_leave) / mu2[from];
  q2[to] = (mu[to] * q[to] + m * q_arrive) / mu2[to];
  if (cl) cl[to] = (mu[to] * cl[to] + m * cl[from]) / mu2[to];   /* (the air takes its cloud water along) */
  const double Ea = column_energy(L, mu2, pib2, th2, q2, par);
  th2[to] += (Eb - Ea) / (mu2[to] * pib2[to]);
  for (int k = 0; k < L; k++) {
    th[k] = th2[k];
    mu[k] = mu2[k];
    pib[k] = pib2[k];
    q[k] = q2[k];
  }
}

/* The dry crossing of every interface theta*_j by the heating dth (air.cu:744-850): wxo[j] the mass moved up across
 * interface j. Dry air (q = 0): the edge-vapor and freeze-drying options carry no vapor and move the same mass. The
 * mixed bottom layer (bl_mix) needs L >= 3 (its top edge is then an interface). dtop/dbot (optional): the heating at
 * each layer's top and bottom edge when it differs from the layer's mean (the k-tables' edge sub-levels). */
static double edge_q(int L, int k, int top, const double* mu, const double* th, const double* q, double m,
                     const double* par, double eps, int mix);
static double qsat_x(double T, double p, double eps, double ice, int mix);
static double fpow_(double x, double y, int mix);

static void crossing(int L, const double* par, double* mu, double* th, double* q, const double* dth, double dth_up,
                     double* pib, double* wxo, const double* dtop, const double* dbot, double* cl, double bl_room,
                     double eps, int mix, double* gross) {
  const double floor_ = par[P_FLOOR];
  for (int j = 0; j + 1 < L; j++) wxo[j] = 0;
  if (par[P_THETA_STAR] <= 0) return;
  double dmj[OG_LMAX], thcj[OG_LMAX], efa[OG_LMAX], efb[OG_LMAX];
  int got[OG_LMAX] = {0};
#define DENS(k, fa, fb)                                                                         \
  do {                                                                                          \
    if (!got[k]) {                                                                              \
      edge_density(tstar((k), L, par), tstar((k) - 1, L, par), th[k] - dth[k], &efa[k], &efb[k]); \
      got[k] = 1;                                                                               \
    }                                                                                           \
    fa = efa[k];                                                                                \
    fb = efb[k];                                                                                \
  } while (0)
  for (int j = 0; j + 1 < L; j++) {
    const int a = j, b = j + 1;
    const double ts_ = tstar(j, L, par);
    double dm = 0, thc = ts_, thd = ts_;
    const int blm = par[P_BL_MIX] > 0;
    if (gross) gross[2 * j] = gross[2 * j + 1] = 0;
    if (dth[b] > 0 && blm && b == L - 1) {
      if (th[b] >= ts_) {                              /* the mixed bottom layer takes in the air above */
        double fa, fb;
        DENS(a, fa, fb);
        dm = -fmin(fmin(mu[a] * dth[b] * fa, fmax(mu[a] - par[P_MU_BL], 0.0) * par[P_BL_CAP]), bl_room);
        thd = ts_ - 0.5 * dth[b];
      }
    } else if (blm && b < L - 1 && th[b] >= ts_) {
      dm = mu[b] - floor_;
      thc = th[b];
    } else if ((dtop ? dtop[b] : dth[b]) > 0) {
      const double eh = dtop ? dtop[b] : dth[b];     /* the heating at b's top edge */
      if (th[b] >= ts_) {
        dm = mu[b] - floor_;
        thc = th[b];
      } else if (b == L - 1) {
        dm = mu[b] * eh / (2 * (ts_ - (th[b] - dth[b])));
      } else {
        double fa, fb;
        DENS(b, fa, fb);
        dm = mu[b] * eh * fb;
      }
    }
    const double da = dbot ? dbot[a] : (a == 0 ? dth_up : dth[a]);
    if (blm && a >= 1 && th[a] <= ts_ && dm <= 0) {
      const double down = mu[a] - floor_;
      thd = th[a];
      dm -= down;
      if (gross) gross[2 * j + 1] = down;
    } else if (da < 0) {
      double down;
      if (th[a] <= ts_) {
        down = mu[a] - floor_;
        thd = th[a];
      } else if (a == 0) {
        down = mu[a] * (-da) / (2 * ((th[a] - dth[a]) - ts_));
      } else {
        double fa, fb;
        DENS(a, fa, fb);
        down = mu[a] * (-da) * fa;
      }
      dm -= down;
      if (gross) gross[2 * j + 1] = down;
    }
    if (gross) gross[2 * j] = dm + gross[2 * j + 1];   /* (the up and down parts before netting and limiting) */
    dmj[j] = dm;
    thcj[j] = dm > 0 ? thc : thd;
  }
#undef DENS
  for (int j = L - 2; j >= 0; j--) {
    const int a = j, b = j + 1;
    double dm = dmj[j];
    const double maxup = mu[b] - floor_, maxdn = mu[a] - floor_;
    if (dm > maxup) dm = maxup > 0 ? maxup : 0.0;
    if (-dm > maxdn) dm = maxdn > 0 ? -maxdn : 0.0;
    /* edge vapor: the crossing air carries the vapor at the edge it leaves through; freeze-drying: air rising into the
     * stratosphere past the cold point at theta* holds no more than saturation there, the rest staying below */
    const int ev = par[P_EDGE_VAP] > 0;
    if (dm > 0 && a == (int)par[P_EDGE_K] - 1 && L >= 3 && par[P_FREEZE_DRY] > 0) {
      double pi_ = 0;
      for (int k = 0; k <= a; k++) pi_ += par[P_GRAV] * mu[k];
      const double Ti = tstar(a, L, par) * fpow_(pi_ / par[P_P0], par[P_KAPPA], mix);
      const double qx = fmin(ev ? edge_q(L, b, 1, mu, th, q, dm, par, eps, mix) : q[b],
                             Ti < 160 ? 0.0 : qsat_x(Ti, pi_, eps, par[P_ICE_AIR], mix));
      const double ma = mu[a], mb = mu[b], ca = cl ? cl[a] : 0.0, cb = cl ? cl[b] : 0.0;
      move_mass(L, par, mu, th, q, pib, b, a, dm, thcj[j], thcj[j], qx, qx, cl);
      if (cl) {
        cl[a] = ma * ca / mu[a];
        cl[b] = mb * cb / mu[b];
      }
    } else if (dm > 0) {
      const double qx = ev ? edge_q(L, b, 1, mu, th, q, dm, par, eps, mix) : q[b];
      move_mass(L, par, mu, th, q, pib, b, a, dm, thcj[j], thcj[j], qx, qx, cl);
    } else if (dm < 0) {
      const double qx = ev ? edge_q(L, a, 0, mu, th, q, -dm, par, eps, mix) : q[a];
      move_mass(L, par, mu, th, q, pib, a, b, -dm, thcj[j], thcj[j], qx, qx, cl);
    }
    wxo[j] = dm;
  }
  const double tiny = floor_ * 1.000001;
  for (int j = L - 2; j >= 0; j--) {
    const int a = j, b = j + 1;
    const double ts_ = tstar(j, L, par);
    if (mu[b] <= tiny && th[b] > ts_) {
      th[a] += mu[b] * pib[b] * (th[b] - ts_) / (mu[a] * pib[a]);
      th[b] = ts_;
    }
    if (mu[a] <= tiny && th[a] < ts_) {
      th[b] -= mu[a] * pib[a] * (ts_ - th[a]) / (mu[b] * pib[b]);
      th[a] = ts_;
    }
  }
}

/* hs_column (air.cu:2065-2109): Newtonian relaxation toward Held & Suarez's equilibrium, then the crossing. */
void og_hs_column(int N, int L, const double* par, const double* hs, const double* lat, const double* mu_in,
                  double* theta, double dt, double* wx, double* qnet, double* stored) {
  const double cp = par[P_CP], kappa = par[P_KAPPA], p0 = par[P_P0];
#pragma omp parallel for schedule(static)
  for (int i = 0; i < N; i++) {
    double mu[OG_LMAX], th[OG_LMAX], q[OG_LMAX], pb[OG_LMAX], pib[OG_LMAX], dth[OG_LMAX], wxo[OG_LMAX];
    for (int k = 0; k < L; k++) {
      mu[k] = mu_in[(size_t)k * N + i];
      th[k] = theta[(size_t)k * N + i];
      q[k] = 0;
    }
    layer_exner(L, mu, par, pb, pib);
    double E0 = 0, Qs = 0;
    for (int k = 0; k < L; k++) E0 += mu[k] * pib[k] * th[k];
    const double s2 = sin(lat[i]) * sin(lat[i]), c2 = 1 - s2;
    const double ps = pb[L - 1];
    double pt = 0;
    for (int k = 0; k < L; k++) {
      const double pbar = pow(pib[k] / cp, 1 / kappa) * p0;
      double Teq = (315 - 60 * s2 - 10 * log(pbar / p0) * c2) * pow(pbar / p0, kappa);
      if (Teq < 200) Teq = 200;
      const double theq = Teq * cp / pib[k];
      const double st = pt / ps, sb = pb[k] / ps;
      const double lo = st > 0.7 ? st : 0.7;
      const double f = sb > 0.7 ? ((sb - 0.7) * (sb - 0.7) - (lo - 0.7) * (lo - 0.7)) / (2 * 0.3 * (sb - st)) : 0.0;
      const double kt = hs[0] + (hs[1] - hs[0]) * f * c2 * c2;
      const double d = (theq - th[k]) * (1 - exp(-kt * dt));
      dth[k] = d;
      th[k] += d;
      Qs += mu[k] * pib[k] * d / dt;
      pt = pb[k];
    }
    double dth_up = dth[0];
    if (L == 2) {
      const double pi_ = pb[0], si = pi_ / ps;
      double Ti = (315 - 60 * s2 - 10 * log(pi_ / p0) * c2) * pow(pi_ / p0, kappa);
      if (Ti < 200) Ti = 200;
      const double thi = Ti * pow(p0 / pi_, kappa);
      const double ki = hs[0] + (hs[1] - hs[0]) * (si > 0.7 ? (si - 0.7) / 0.3 : 0.0) * c2 * c2;
      dth_up = (thi - par[P_THETA_STAR]) * (1 - exp(-ki * dt));
    }
    crossing(L, par, mu, th, q, dth, dth_up, pib, wxo, NULL, NULL, NULL, 1e30, 0.622, 0, NULL);   /* (dry: no vapor moves) */
    for (int j = 0; j + 1 < L; j++) wx[(size_t)j * N + i] = wxo[j];
    double E1 = 0;
    for (int k = 0; k < L; k++) {
      theta[(size_t)k * N + i] = th[k];
      E1 += mu[k] * pib[k] * th[k];
    }
    qnet[i] = Qs;
    stored[i] = E1 - E0;
  }
}

/* Tetens over water (air.cu:157-159), with the air's eps. */
static double qsat(double T, double p, double eps) { return eps * 611.0 * exp(17.27 * (T - 273.15) / (T - 35.85)) / p; }

/* surface_layer (air.cu:2015-2061): Louis's bulk exchange with Charnock's sea roughness, four fixed-point passes. */
void og_surface_layer(int N, int L, const double* par, const og_air* air, const double* uabs, const double* mu_in,
                      const double* theta, const double* qv, const double* Ts, const double* ice, const double* land,
                      double* cd, double* ch, const double* lc) {
  const double grav = par[P_GRAV], p0 = par[P_P0], kappa = par[P_KAPPA], rd = air->rd, vf = air->vf;
  const double moist = par[P_MOIST], q_sea = par[P_Q_SEA], gust = par[P_BULK_GUST];
#pragma omp parallel for schedule(static)
  for (int i = 0; i < N; i++) {
    const int bot = L - 1;
    double ps = 0;
    for (int k = 0; k < L; k++) ps += grav * mu_in[(size_t)k * N + i];
    const double mb = mu_in[(size_t)bot * N + i], th = theta[(size_t)bot * N + i];
    const double Ta = th * pow(ps / p0, kappa), T0 = Ts[i];
    const double rho = ps / (rd * Ta);
    const double qa = qv ? qv[(size_t)bot * N + i] : 0.0;
    const int isl = land && land[i] > 0.5;
    const double q0 = qv && moist > 0 && !isl ? (q_sea > 0 ? q_sea : 1.0) * qsat(T0, ps, air->eps) : qa;
    const double fice = qv && moist > 0 && ice && !isl ? fmin(1.0, ice[i] / par[P_RHO_ICE] / par[P_ICE_COVER]) : 0.0;
    const double depth = rd * Ta / grav * log(ps / fmax(ps - grav * mb, 1.0));
    const double zr = fmax(10.0, 0.1 * depth);
    const double dthv = (Ta - T0) + vf * Ta * (qa - q0);
    const double K = 0.4, nu = 1.5e-5, b = 5, c = 5;
    double U = sqrt(uabs[i] * uabs[i] + gust * gust), Cd = 1e-3, Ch = 1e-3;
    for (int it = 0; it < 4; it++) {
      const double us = fmax(sqrt(Cd) * U, 1e-3);
      double z0, zh;
      if (isl) {
        z0 = lc ? lc[i] : par[P_Z0_LAND];
        zh = 0.1 * z0;
      } else {
        const double zs = 0.011 * us * us / grav + 0.11 * nu / us, zsh = 0.40 * nu / us;
        z0 = (1 - fice) * zs + fice * 1e-3;
        zh = (1 - fice) * zsh + fice * 1e-4;
      }
      const double lm = log(zr / z0), lh = log(zr / zh);
      const double CdN = K * K / (lm * lm), ChN = K * K / (lm * lh);
      const double Ri = grav * zr * dthv / (Ta * U * U);
      double fm, fh;
      if (Ri > 0) {
        const double s = sqrt(1 + c * Ri);
        fm = 1 / (1 + 2 * b * Ri / s);
        fh = 1 / (1 + 3 * b * Ri * s);
      } else {
        const double d = 1 + 3 * b * c * CdN * sqrt(zr / z0 * -Ri);
        fm = 1 - 2 * b * Ri / d;
        fh = 1 - 3 * b * Ri / d;
      }
      Cd = CdN * fm;
      Ch = ChN * fh;
      const double B = -Ch * U * dthv;
      const double ws = B > 0 ? cbrt(grav / Ta * B * depth) : 0.0;
      U = sqrt(uabs[i] * uabs[i] + gust * gust + 1.44 * ws * ws);
    }
    cd[i] = rho * Cd * U;
    ch[i] = rho * Ch * U;
  }
}

void og_cell_speed(int N, const double* u, const int32_t* cellEdge, const double* le, const double* de,
                   const double* areaCell, double* out) {
#pragma omp parallel for schedule(static)
  for (int i = 0; i < N; i++) {
    double ke = 0;
    for (int m = 0; m < 6; m++) {
      const int e = cellEdge[i * 6 + m];
      if (e < 0) break;
      ke += le[e] * de[e] * 0.25 * u[e] * u[e];
    }
    out[i] = sqrt(2 * ke / areaCell[i]);
  }
}

void og_cell_speed_perot(int N, const double* u, const int32_t* cellEdge, const double* op, double* out) {
#pragma omp parallel for schedule(static)
  for (int i = 0; i < N; i++) {
    double ue = 0, un = 0;
    for (int m = 0; m < 6; m++) {
      const int e = cellEdge[i * 6 + m];
      if (e < 0) break;
      ue += op[(i * 6 + m) * 2] * u[e];
      un += op[(i * 6 + m) * 2 + 1] * u[e];
    }
    out[i] = sqrt(ue * ue + un * un);
  }
}

void og_gas_src(int N, int L, int G, const double* src, const double* h, double dt, double* x) {
  for (int g = 0; g < G; g++)
    for (int i = 0; i < N; i++) x[((size_t)g * L + L - 1) * N + i] += src[(size_t)g * N + i] * dt / h[(size_t)(L - 1) * N + i];
}

/* ---- air_column (air.cu:876-1972): the dry column, and the moist one with clouds and land, without the mixed layer ----
 * The sun (mode 0: the diurnal and seasonal sun at t; mode 1: insol), the shortwave split (the k-tables' swf, or by
 * the s^4 or tau profile, or by mass), the gray longwave (the vapor's part from the column's vapor with lw_w0; or the
 * k-tables' radf), the slab surface (or, moist, the sea-ice skin; over land the ground with its soil bucket, ponds,
 * snow, firn and skin) under the bulk or surface-layer exchange, the evaporation, the heating as theta, the
 * condensation (with clouds the cloud scheme), the crossing with the k-tables' edge heating, and the moist
 * convection's parcel plume. `mix` selects the reference's mixed build (air.cu:24-36): 1 - e^-x, the exponentials and
 * powers of the saturation and condensation arithmetic, and that arithmetic itself in float32 (`fast`). */
enum { P_S0 = 0, P_ASW = 1, P_TAU0 = 2, P_MU_REF = 3, P_EPS_S = 4, P_H_SENS = 5, P_U0 = 6, P_ALB = 7, P_C_SLAB = 8,
       P_SIGMA = 12, P_TILT = 17, P_DAY = 19, P_TAU_LIN = 20, P_LF = 24, P_TF = 25, P_K_EVAP = 26, P_TAU_CONV = 27,
       P_ALB_ICE = 28, P_K_ICE = 31, P_RV = 32, P_SAT_LAYER = 33, P_SW_TAU = 34, P_ALB_ICE_MELT = 41, P_ALB_RAMP = 45,
       P_LW_W0 = 48, P_BL_BOT = 54, P_THETA_RISE = 69, P_BULK_C = 74, P_BULK_Z = 76, P_BULK_W = 77, P_EDGE_HEAT = 84,
       P_SFC = 90, P_DRY_EDGE = 99, P_FIRN = 120, P_RHO_FIRN = 122, P_C_ICE = 123,
       P_CLOUD = 55, P_RH_CRIT = 56, P_RE_LIQ = 57, P_RE_ICE = 58, P_CLOUD_G = 59, P_K_LW_LIQ = 60, P_K_LW_ICE = 61,
       P_C_AUTO = 62, P_Q_AUTO = 63, P_V_ICE = 64, P_TAU_ERODE = 66, P_K_REVAP = 67, P_MU_CLOUD = 68, P_L_ICE = 79,
       P_CONV_REVAP = 87, P_DOWNDRAFT = 92, P_REVAP = 103, P_REVAP_RHS = 104, P_REVAP_AREA = 106, P_LIQ_COLD = 128,
       P_CW_QS = 129, P_ALB_DRY = 35, P_ALB_WET = 36, P_SOIL_BETA = 37, P_K_INF = 38, P_POND_COVER = 39,
       P_ALB_SNOW = 40, P_ALB_SNOW_MELT = 42, P_ALB_GLACIER = 43, P_SNOW_COVER = 44, P_LAND_EXCH = 46, P_C_DRY = 49,
       P_C_WET = 50, P_LAND_RS = 52, P_C_DEEP = 94, P_RS_MAX = 95, P_RS_SUN = 96, P_REVAP_RHL = 105, P_BARE_FX = 115,
       P_K_FIRN = 121, P_MU_NIGHT = 71, P_BL_ENT = 72, P_BL_REACH = 98, P_BL_LCL = 100, P_ENT_LCL = 130 };

static double omexp(double x, int mix) { return mix ? (double)-expm1f(-(float)x) : 1 - exp(-(x)); }
static double fpow_(double x, double y, int mix) { return mix ? (double)powf((float)x, (float)y) : pow(x, y); }

/* The mass share m's theta in a layer of shape s as a fraction of the way between its edges, ln(1 + m (e^s - 1)) / s
 * (air.cu:233-244; es = e^s - 1). The mixed build in float32 with asymptotes beyond |s| = 30; the shape s itself
 * comes from the FP64 edge_shape in both forms. */
static double spread_of(double m, double s, double es, int mix) {
  if (mix) {
    const float mf = (float)m, sf = (float)s;
    if (fabsf(sf) < 1e-6f) return mf;
    if (sf > 30.f) return 1.f + logf(mf) / sf;
    if (sf < -30.f) return log1pf(-mf) / sf;
    return log1pf(mf * expm1f(sf)) / sf;
  }
  return fabs(s) < 1e-6 ? m : log1p(m * es) / s;
}

/* tau_p (air.cu:853-856): the gray optical depth down to pressure p of a column with surface pressure ps. */
static double tau_p(double p, double ps, const double* par) {
  const double s = p / ps, s2 = s * s;
  return par[P_TAU0] * (par[P_TAU_LIN] * s + (1 - par[P_TAU_LIN]) * s2 * s2);
}

/* qsat and qsat_a (air.cu:156-171): Tetens over water; with ice_air, below freezing the blend with Tetens over ice
 * (Murray 1967) by the liquid share (1 at -20 C ... 0). In the mixed build in float32 throughout. */
static double qsat_x(double T, double p, double eps, double ice, int mix) {
  if (mix) {
    const float t = (float)T, pp = (float)p, e = (float)eps;
    const float w = e * 611.0f * expf(17.27f * (t - 273.15f) / (t - 35.85f)) / pp;
    if (!(ice > 0) || t >= 273.15f) return w;
    const float l = fminf(1.0f, fmaxf(0.0f, (t - 253.15f) / 20.0f));
    const float si = e * 611.0f * expf(21.875f * (t - 273.16f) / (t - 7.66f)) / pp;
    return l * w + (1 - l) * si;
  }
  const double w = eps * 611.0 * exp(17.27 * (T - 273.15) / (T - 35.85)) / p;
  if (!(ice > 0) || T >= 273.15) return w;
  const double l = fmin(1.0, fmax(0.0, (T - 253.15) / 20.0));
  const double si = eps * 611.0 * exp(21.875 * (T - 273.16) / (T - 7.66)) / p;
  return l * w + (1 - l) * si;
}

/* condense (air.cu:190-198): saturation adjustment at Exner pib and pressure p in one Newton step. */
static double condense(double* th, double* q, double pib, double p, const double* par, double eps, int mix) {
  const double Lv = par[P_LV], cp = par[P_CP], Rv = par[P_RV], ice = par[P_ICE_AIR];
  double dq;
  if (mix) {
    const float T = (float)pib * (float)*th / (float)cp, qs = (float)qsat_x(T, (float)p, eps, ice, 1);
    if (!(*q > qs)) return 0;
    dq = (float)(*q - qs) / (1.0f + (float)Lv * (float)Lv * qs / ((float)cp * (float)Rv * T * T));
  } else {
    const double T = pib * *th / cp, qs = qsat_x(T, p, eps, ice, 0);
    if (!(*q > qs)) return 0;
    dq = (*q - qs) / (1 + Lv * Lv * qs / (cp * Rv * T * T));
  }
  *q -= dq;
  *th += Lv * dq / pib;
  return dq;
}

/* condense_layer (air.cu:275-291): the capacity averaged over the layer's pressure span (4-point Gauss-Legendre,
 * nodes colder than 160 K left out), one Newton step with Clausius-Clapeyron's slope at each node. */
static double condense_layer(double* th, double* q, double pib, double pt, double pb, const double* par, double eps,
                             int mix) {
  static const double x[4] = {-0.8611363115940526, -0.3399810435848563, 0.3399810435848563, 0.8611363115940526};
  static const double wq[4] = {0.3478548451374538, 0.6521451548625461, 0.6521451548625461, 0.3478548451374538};
  const double Lv = par[P_LV], Rv = par[P_RV], ice = par[P_ICE_AIR], p0 = par[P_P0], kappa = par[P_KAPPA];
  double dq;
  if (mix) {
    float qs = 0, dqs = 0;
    for (int j = 0; j < 4; j++) {
      const double p = 0.5 * (pt + pb) + 0.5 * (pb - pt) * x[j];
      const float ex = powf((float)(p / p0), (float)kappa), T = (float)*th * ex;
      if (T < 160) continue;
      const float s = (float)qsat_x(T, (float)p, eps, ice, 1);
      qs += 0.5f * (float)wq[j] * s;
      dqs += 0.5f * (float)wq[j] * s * (float)Lv / ((float)Rv * T * T) * ex;
    }
    if (!(*q > qs)) return 0;
    dq = (float)(*q - qs) / (1.0f + (float)Lv / (float)pib * dqs);
  } else {
    double qs = 0, dqs = 0;
    for (int j = 0; j < 4; j++) {
      const double p = 0.5 * (pt + pb) + 0.5 * (pb - pt) * x[j];
      const double ex = pow(p / p0, kappa), T = *th * ex;
      if (T < 160) continue;
      const double s = qsat_x(T, p, eps, ice, 0);
      qs += 0.5 * wq[j] * s;
      dqs += 0.5 * wq[j] * s * Lv / (Rv * T * T) * ex;
    }
    if (!(*q > qs)) return 0;
    dq = (*q - qs) / (1 + Lv / pib * dqs);
  }
  *q -= dq;
  *th += Lv * dq / pib;
  return dq;
}

/* condense_spread (air.cu:292-315): a middle layer's saturation adjustment with its theta spread between its edges a
 * and b as edge_shape has it, the whole spread shifting with the mean; one Newton step. */
static double condense_spread(double a, double b, double* th, double* q, double pib, double pt, double pb,
                              const double* par, double eps, int mix) {
  static const double x[4] = {-0.8611363115940526, -0.3399810435848563, 0.3399810435848563, 0.8611363115940526};
  static const double wq[4] = {0.3478548451374538, 0.6521451548625461, 0.6521451548625461, 0.3478548451374538};
  const double Lv = par[P_LV], Rv = par[P_RV], ice = par[P_ICE_AIR], p0 = par[P_P0], kappa = par[P_KAPPA];
  const double s = edge_shape(a, b, *th), es = mix ? 0.0 : expm1(s);
  double dq;
  if (mix) {
    float qs = 0, dqs = 0;
    for (int j = 0; j < 4; j++) {
      const double p = 0.5 * (pt + pb) + 0.5 * (pb - pt) * x[j], m = (pb - p) / (pb - pt);
      const double tn = a + (b - a) * spread_of(m, s, es, 1);
      const float ex = powf((float)(p / p0), (float)kappa), T = (float)tn * ex;
      if (T < 160) continue;
      const float v = (float)qsat_x(T, (float)p, eps, ice, 1);
      qs += 0.5f * (float)wq[j] * v;
      dqs += 0.5f * (float)wq[j] * v * (float)Lv / ((float)Rv * T * T) * ex;
    }
    if (!(*q > qs)) return 0;
    dq = (float)(*q - qs) / (1.0f + (float)Lv / (float)pib * dqs);
  } else {
    double qs = 0, dqs = 0;
    for (int j = 0; j < 4; j++) {
      const double p = 0.5 * (pt + pb) + 0.5 * (pb - pt) * x[j], m = (pb - p) / (pb - pt);
      const double tn = a + (b - a) * spread_of(m, s, es, 0);
      const double ex = pow(p / p0, kappa), T = tn * ex;
      if (T < 160) continue;
      const double v = qsat_x(T, p, eps, ice, 0);
      qs += 0.5 * wq[j] * v;
      dqs += 0.5 * wq[j] * v * Lv / (Rv * T * T) * ex;
    }
    if (!(*q > qs)) return 0;
    dq = (*q - qs) / (1 + Lv / pib * dqs);
  }
  *q -= dq;
  *th += Lv * dq / pib;
  return dq;
}

/* condense_k (air.cu:319-331): the layer's capacity (sat_layer; sat_prof: a middle layer below edge_k's tropopause
 * with its theta spread between its edges; sat_base: the third layer at its base), else v1's rule at the pressure
 * whose Exner function is the layer's. */
static double condense_k(int L, int k, const double* mu, double* th, double* q, const double* pib, const double* par,
                         double eps, int mix) {
  if (par[P_SAT_LAYER] > 0) {
    double pt = 0;
    for (int j = 0; j < k; j++) pt += par[P_GRAV] * mu[j];
    if (par[P_SAT_PROF] > 0 && k >= (int)par[P_EDGE_K] && k < L - 1 && th[k] < tstar(k - 1, L, par))
      return condense_spread(tstar(k, L, par), tstar(k - 1, L, par), &th[k], &q[k], pib[k], pt,
                             pt + par[P_GRAV] * mu[k], par, eps, mix);
    if (par[P_SAT_BASE] > 0 && k == 2) {
      const double pb = pt + par[P_GRAV] * mu[k];
      return condense(&th[k], &q[k], par[P_CP] * fpow_(pb / par[P_P0], par[P_KAPPA], mix), pb, par, eps, mix);
    }
    return condense_layer(&th[k], &q[k], pib[k], pt, pt + par[P_GRAV] * mu[k], par, eps, mix);
  }
  const double p = fpow_(pib[k] / par[P_CP], 1 / par[P_KAPPA], mix) * par[P_P0];
  return condense(&th[k], &q[k], pib[k], p, par, eps, mix);
}

/* parcel (air.cu:468-473): air (th, q) lifted to pressure p, condensing what it cannot hold. */
static void parcel(double th, double q, double p, const double* par, double eps, int mix, double* thp, double* qp) {
  *thp = th;
  *qp = q;
  const double ex = fpow_(p / par[P_P0], par[P_KAPPA], mix);
  condense(thp, qp, par[P_CP] * ex, p, par, eps, mix);
}

/* layer_cap (air.cu:334-359, the layer's theta uniform): the layer's vapor capacity over its pressure span and its
 * change with theta, as condense_layer takes them. */
static void layer_cap(int L, int k, const double* mu, double th, const double* par, double eps, int mix, double* qs,
                      double* dqs) {
  static const double x[4] = {-0.8611363115940526, -0.3399810435848563, 0.3399810435848563, 0.8611363115940526};
  static const double wq[4] = {0.3478548451374538, 0.6521451548625461, 0.6521451548625461, 0.3478548451374538};
  const double Lv = par[P_LV], Rv = par[P_RV], ice = par[P_ICE_AIR], p0 = par[P_P0], kappa = par[P_KAPPA];
  double pt = 0;
  for (int j = 0; j < k; j++) pt += par[P_GRAV] * mu[j];
  const double pb = pt + par[P_GRAV] * mu[k];
  /* (sat_prof: a middle layer below the tropopause with its theta spread between its edges) */
  const int spread = par[P_SAT_PROF] > 0 && k >= (int)par[P_EDGE_K] && k < L - 1 && th < tstar(k - 1, L, par);
  const double ea = spread ? tstar(k, L, par) : 0.0, eb = spread ? tstar(k - 1, L, par) : 0.0;
  const double s = spread ? edge_shape(ea, eb, th) : 0.0, es = spread && !mix ? expm1(s) : 0.0;
  if (mix) {
    float qf = 0, dqf = 0;
    for (int j = 0; j < 4; j++) {
      const double p = 0.5 * (pt + pb) + 0.5 * (pb - pt) * x[j], m = (pb - p) / (pb - pt);
      const double tn = !spread ? th : ea + (eb - ea) * spread_of(m, s, es, 1);
      const float ex = powf((float)(p / p0), (float)kappa), T = (float)tn * ex;
      if (T < 160) continue;
      const float v = (float)qsat_x(T, (float)p, eps, ice, 1);
      qf += 0.5f * (float)wq[j] * v;
      dqf += 0.5f * (float)wq[j] * v * (float)Lv / ((float)Rv * T * T) * ex;
    }
    *qs = qf;
    *dqs = dqf;
  } else {
    double qf = 0, dqf = 0;
    for (int j = 0; j < 4; j++) {
      const double p = 0.5 * (pt + pb) + 0.5 * (pb - pt) * x[j], m = (pb - p) / (pb - pt);
      const double tn = !spread ? th : ea + (eb - ea) * spread_of(m, s, es, 0);
      const double ex = pow(p / p0, kappa), T = tn * ex;
      if (T < 160) continue;
      const double v = qsat_x(T, p, eps, ice, 0);
      qf += 0.5 * wq[j] * v;
      dqf += 0.5 * wq[j] * v * Lv / (Rv * T * T) * ex;
    }
    *qs = qf;
    *dqs = dqf;
  }
}

/* edge_q (air.cu:361-378): the vapor of layer k's air at its top edge (top) or bottom edge, its relative humidity
 * against the layer's capacity held uniform through the layer (a spread layer's edge at its theta*); layers above
 * edge_k (with L >= 3) and, with bl_q, the mixed bottom layer keep their vapor even. Air of mass m leaving through
 * that edge takes the edge's vapor plus m's share of the way to the mean, at most the layer's whole vapor. */
static double edge_q(int L, int k, int top, const double* mu, const double* th, const double* q, double m,
                     const double* par, double eps, int mix) {
  if (!(q[k] > 0) || (k < (int)par[P_EDGE_K] && L >= 3) || (par[P_BL_Q] > 0 && k == L - 1 && k > 0)) return q[k];
  double qs, dqs;
  layer_cap(L, k, mu, th[k], par, eps, mix, &qs, &dqs);
  if (!(qs > 0)) return q[k];
  double pt = 0;
  for (int j = 0; j < k; j++) pt += par[P_GRAV] * mu[j];
  const double p = top ? pt : pt + par[P_GRAV] * mu[k];
  if (!(p > 0)) return q[k];
  const int spread = par[P_SAT_PROF] > 0 && k >= (int)par[P_EDGE_K] && k < L - 1 && th[k] < tstar(k - 1, L, par);
  const double te = spread ? (top ? tstar(k - 1, L, par) : tstar(k, L, par)) : th[k];
  const double T = te * fpow_(p / par[P_P0], par[P_KAPPA], mix);
  const double qe = T < 160 ? 0.0 : q[k] * qsat_x(T, p, eps, par[P_ICE_AIR], mix) / qs;
  const double f = m > 0 ? fmin(1.0, m / mu[k]) : 0.0;
  return m > 0 ? fmin(qe + (q[k] - qe) * f, mu[k] * q[k] / m) : qe;
}

/* cloud_pdf (air.cu:383-388): the water spread evenly over +-(1 - rh_crit) qs about its mean qt; its saturated share
 * f and the mean condensate l there. */
static void cloud_pdf(double qt, double qs, const double* par, double* l, double* f) {
  const double b = (1 - par[P_RH_CRIT]) * qs, s = qt - qs;
  if (s <= -b) {
    *l = 0;
    *f = 0;
  } else if (s >= b) {
    *l = s;
    *f = 1;
  } else {
    *f = (s + b) / (2 * b);
    *l = (s + b) * *f / 2;
  }
}

/* cover_of (air.cu:393-396): the cover of cloud water cw spread at a cloud's typical water content by phase. */
static double cover_of(double cw, double mu, double liq, const double* par, double qs) {
  const double s = par[P_CW_QS] > 0 && qs > 0 ? fmin(1.0, par[P_CW_QS] * qs / par[P_Q_AUTO]) : 1.0;
  return fmin(1.0, cw / (s * (liq * par[P_Q_AUTO] + (1 - liq) * par[P_L_ICE])) * fmax(1.0, mu / par[P_MU_CLOUD]));
}

/* mixed_cloud (air.cu:397-426, bl_q): the mixed bottom layer's vapor is even through its depth, so it is nearest
 * saturation at its top: its cloud statistics at each of layer_cap's four levels, l the mean condensate, f the cover
 * where it is most, dqf the condensate's fall per kelvin of warming. ml_clear (unless pdf): only air actually
 * saturated condenses there. Returns 0 for any other layer (or without bl_q). */
static int mixed_cloud(int L, int k, const double* mu, double th, double qt, const double* par, double eps, int mix,
                       double* l, double* f, double* dqf, int pdf) {
  static const double x[4] = {-0.8611363115940526, -0.3399810435848563, 0.3399810435848563, 0.8611363115940526};
  static const double wq[4] = {0.3478548451374538, 0.6521451548625461, 0.6521451548625461, 0.3478548451374538};
  if (!(par[P_BL_Q] > 0 && k == L - 1 && k > 0)) return 0;
  double pt = 0;
  for (int j = 0; j < k; j++) pt += par[P_GRAV] * mu[j];
  const double pb = pt + par[P_GRAV] * mu[k];
  *l = 0;
  *f = 0;
  *dqf = 0;
  for (int j = 0; j < 4; j++) {
    const double p = 0.5 * (pt + pb) + 0.5 * (pb - pt) * x[j];
    const double ex = fpow_(p / par[P_P0], par[P_KAPPA], mix), T = th * ex;
    if (T < 160) continue;
    const double sq = qsat_x(T, p, eps, par[P_ICE_AIR], mix);
    double lj, fj;
    if (par[P_ML_CLEAR] > 0 && !pdf) {
      lj = fmax(0.0, qt - sq);
      fj = qt > sq ? 1.0 : 0.0;
    } else {
      cloud_pdf(qt, sq, par, &lj, &fj);
    }
    *l += 0.5 * wq[j] * lj;
    *f = fmax(*f, fj);
    *dqf += 0.5 * wq[j] * fj * sq * par[P_LV] / (par[P_RV] * T * T) * ex;
  }
  return 1;
}

/* cloud_adjust (air.cu:438-450): condense up to the statistics' condensate with the latent heat's feedback on the
 * capacity; a cloud holding more erodes over tau_erode. */
static int mixed_cloud(int L, int k, const double* mu, double th, double qt, const double* par, double eps, int mix,
                       double* l, double* f, double* dqf, int pdf);

static double cloud_adjust(int L, int k, const double* mu, double* th, double* q, double* cw, const double* pib,
                           double dt, const double* par, double eps, int mix) {
  double qs, dqs, l, f;
  if (mixed_cloud(L, k, mu, th[k], q[k] + cw[k], par, eps, mix, &l, &f, &dqs, 0)) {
    f = 1;                                           /* (dqs then already carries each level's share) */
  } else {
    layer_cap(L, k, mu, th[k], par, eps, mix, &qs, &dqs);
    if (!(qs > 0)) return 0;
    cloud_pdf(q[k] + cw[k], qs, par, &l, &f);
  }
  const double dl = l > cw[k] ? (l - cw[k]) / (1 + par[P_LV] / pib[k] * dqs * f)
                              : -(cw[k] - l) * omexp(dt / par[P_TAU_ERODE], mix);
  cw[k] += dl;
  q[k] -= dl;
  th[k] += par[P_LV] * dl / pib[k];
  return dl;
}

/* base_T (air.cu:476-482): the temperature at layer k's lower edge: its theta there, or for a spread middle layer
 * (sat_prof) its lower theta*. */
static double base_T(int L, int k, const double* mu, const double* th, const double* par, int mix) {
  double pt = 0;
  for (int j = 0; j < k; j++) pt += par[P_GRAV] * mu[j];
  const double pb = pt + par[P_GRAV] * mu[k];
  const int spread = par[P_SAT_PROF] > 0 && k >= (int)par[P_EDGE_K] && k < L - 1 && th[k] < tstar(k - 1, L, par);
  return (spread ? tstar(k, L, par) : th[k]) * fpow_(pb / par[P_P0], par[P_KAPPA], mix);
}

/* lcl_mass (air.cu:1413-1421): the mixed bottom layer's mass up to the level where its air, lifted
 * dry-adiabatically, saturates (40 bisections in pressure); 0 if saturated at the ground, 1e30 if not even at a tenth
 * of the surface pressure. */
static double lcl_mass(int L, const double* mu, const double* th, const double* q, const double* par, double eps,
                       int mix) {
  const int bot = L - 1;
  double ps_ = 0;
  for (int k = 0; k < L; k++) ps_ += par[P_GRAV] * mu[k];
#define SAT(p_) (qsat_x(th[bot] * fpow_((p_) / par[P_P0], par[P_KAPPA], mix), (p_), eps, par[P_ICE_AIR], mix) <= q[bot])
  if (SAT(ps_)) return 0;
  if (!SAT(0.1 * ps_)) return 1e30;
  double lo = 0.1 * ps_, hi = ps_;
  for (int it = 0; it < 40; it++) {
    const double m_ = 0.5 * (lo + hi);
    if (SAT(m_)) lo = m_;
    else hi = m_;
  }
#undef SAT
  return (ps_ - 0.5 * (lo + hi)) / par[P_GRAV];
}

/* The liquid share of condensate at T (air.cu:160): 1 at freezing, 0 at -20 C (float32 in the mixed build). */
static double liq_of(double T, int mix) {
  if (mix) return fminf(1.0f, fmaxf(0.0f, ((float)T - 253.15f) / 20.0f));
  return fmin(1.0, fmax(0.0, (T - 253.15) / 20.0));
}

/* plume_ascent (air.cu:540-666, the deep plume's options plume_ent, ent_cloud, ent_det, ice_air, ent_ifs, ent_loc,
 * det_all, buoy_v; air.cu:564-575, 583-590, 612-624 for the last four): bottom air (th0,
 * q0) lifted from the bottom layer's top (pi_) up through the layers, in M sub-levels each, condensing as it goes and
 * mixing with (plume_ent) and, through the layer above the bottom one, entraining (ent_cloud, its mass growing as
 * exp(ent_cloud z); ent_det: as much detrained) the air it passes, whose theta at mass share x above the layer's
 * lower edge a is a + (b - a) ln(1 + x (e^s - 1)) / s. A: its buoyancy summed by mass up to the highest level it
 * beats (the bottom layer's stretch a triangle); kd, thd, qd: the layer it detrains into, its theta and vapor there;
 * qx: what its mixing gave each layer per kg; G its mass per kg leaving the bottom layer; FZ its frozen condensate
 * (ice_air); lt: its condensate freezing between 0 and -40 C, those levels' pressures and its CAPE, for lightning.
 * ent_ifs: the entraining layers' entrainment and detrainment the IFS's, 1.75e-3 /m (1.3 - RH) (q_s / q_s at cloud
 * base)^3 and 0.75e-4 /m, in sub-levels of ~25 hPa; det_all: every tropospheric layer it crosses entraining;
 * ent_loc: the air taken in at the layer's relative humidity level by level; buoy_v: buoyancy in virtual theta. */
static void plume_ascent(int L, const double* mu, const double* th, double th0, double q0, double pi_,
                         const double* par, const og_air* air, int mix, double* A, int* kd, double* thd, double* qd,
                         const double* q, double* qx, double* lt, double* G, double* FZ, int shal, double* Gl,
                         double* ptop) {
  const int bot = L - 1, M = bot > 2 ? 3 : 8, cu = par[P_CU_DTH] > 0;
  const double grav = par[P_GRAV], p0 = par[P_P0], kappa = par[P_KAPPA], cp = par[P_CP], Lf = par[P_LF];
  const double Tf = par[P_TF], rd = air->rd, eps = air->eps;
  double Mp = 1, tp = th0, qp = q0, fz = 0;
  /* (cu_dth: the growth it takes in from each layer; the half sub-level a layer's top leaves to the next) */
  double gk[OG_LMAX], dzc = 0;
#define LIFT(p_, ex_)                                                                \
  do {                                                                               \
    const double q_ = qp;                                                            \
    ex_ = fpow_((p_) / p0, kappa, mix);                                              \
    condense(&tp, &qp, cp * ex_, (p_), par, eps, mix);                               \
    if (par[P_ICE_AIR] > 0 && q_ > qp) {                                             \
      const double z_ = (q_ - qp) * (1 - liq_of(tp * ex_, mix));                     \
      tp += Lf * z_ / (cp * ex_);                                                    \
      fz += Mp * z_;                                                                 \
    }                                                                                \
  } while (0)
  double ex0;
  LIFT(pi_, ex0);
  /* buoy_v: buoyancy in virtual theta, theta (1 + 0.608 q), the plume's vapor against its surroundings' */
#define VF(qv_) (par[P_BUOY_V] > 0 ? 1 + 0.608 * (qv_) : 1.0)
  const int ifs = par[P_ENT_IFS] > 0;
  double sum = 0.5 * mu[bot] * (tp * VF(qp) - tstar(bot - 1, L, par) * VF(q[bot - 1])), pt = pi_;
  /* (ent_ifs: the environment's saturation at cloud base, the mixed layer's top) */
  const double qsb = ifs ? qsat_x(th[bot] * fpow_(pi_ / p0, kappa, mix), pi_, eps, par[P_ICE_AIR], mix) : 1.0;
  *A = 0;
  *kd = -1;
  *thd = tp;
  *qd = qp;
  double cape = 0, q0c = -1, q40 = -1, p0c = 0, p40 = 0, qa[OG_LMAX];
  for (int k = 0; k <= bot; k++) {
    qa[k] = 0;
    qx[k] = 0;
    gk[k] = 0;
    if (Gl) Gl[k] = 0;
  }
  for (int k = bot - 1; k >= 0; k--) {
    const double a = tstar(k, L, par), b = k > 0 ? tstar(k - 1, L, par) : 0.0;
    const double s = k > 0 ? edge_shape(a, b, th[k]) : 0.0, es = k > 0 && !mix ? expm1(s) : 0.0;
    /* (ent_ifs: the cloud layer's ascent in sub-levels of ~25 hPa; det_all: the IFS's entrainment and detrainment
     * through every tropospheric layer the plume crosses, below edge_k's stratosphere) */
    /* (cu_dth: and through the cloud layer above the cumulus layer) */
    const int ek = k > 0 && (k == bot - 1 || (cu && ifs && k == bot - 2) ||
                             (par[P_DET_ALL] > 0 && ifs && !shal && k >= (int)par[P_EDGE_K]));
    int Mk = M;
    if (ifs && ek) {
      const int n = (int)ceil(grav * mu[k] / 2500.0);
      Mk = n > M ? n : M;
      if (Mk > 48) Mk = 48;
    }
    const double f = par[P_PLUME_ENT] > 0 ? omexp(par[P_PLUME_ENT] * grav * mu[k] / Mk, mix) : 0.0;
    /* ent_loc: the air taken in at each sub-level holds the layer's vapor at its relative humidity there; qsl the
     * mean of the sub-levels' saturation */
    const int loc = par[P_ENT_LOC] > 0 && par[P_ENT_CLOUD] > 0 && ifs && ek && k >= (int)par[P_EDGE_K] && q[k] > 0;
    double qsl = 0;
    if (loc)
      for (int j = 0; j < Mk; j++) {
        const double x = (j + 0.5) / Mk, pl = pt - x * grav * mu[k];
        const double te = a + (b - a) * spread_of(x, s, es, mix);
        qsl += qsat_x(te * fpow_(pl / p0, kappa, mix), pl, eps, par[P_ICE_AIR], mix) / Mk;
      }
    int last = 0;
    for (int j = 0; j < Mk; j++) {
      const double x = (j + 0.5) / Mk, pl = pt - x * grav * mu[k];
      double ex;
      LIFT(pl, ex);
      const double te = k == 0 ? th[0] : a + (b - a) * spread_of(x, s, es, mix);
      if (par[P_ENT_CLOUD] > 0 && ek) {
        double dz = rd * te * ex / grav * (grav * mu[k] / Mk) / pl;
        if (ifs && j == 0) dz *= 0.5;                 /* (from cloud base to the first sub-level's middle) */
        /* (cu_dth: a layer's first sub-level also takes the rest of the layer below's last one; its last leaves half) */
        if (cu && j == 0) dz += dzc;
        if (cu && j == Mk - 1) dzc = 0.5 * rd * te * ex / grav * (grav * mu[k] / Mk) / pl;
        if (ifs) {
          /* ent_ifs: the IFS's entrainment 1.75e-3 /m (1.3 - RH) (q_s / q_s at cloud base)^3 and mixing detrainment
           * 0.75e-4 /m (Bechtold et al. 2008, 2014): the detrainment an equal-mass exchange, the rest of the
           * entrainment growing the plume */
          const double qse = qsat_x(te * ex, pl, eps, par[P_ICE_AIR], mix), qen = loc && qsl > 0 ? q[k] * qse / qsl : q[k];
          const double rh = fmin(1.0, fmax(0.0, qen / qse));
          const double fs = fmin(1.0, qse / qsb);
          const double eps_ = (shal ? 3.5e-3 : 1.75e-3) * (1.3 - rh) * fs * fs * fs, del_ = fmin(eps_, 0.75e-4);
          const double fe_ = -expm1(-eps_ * dz), fd_ = -expm1(-del_ * dz);
          qa[k] += Mp * fd_ * (qp - qen);
          if (loc) qa[k] -= Mp * expm1((eps_ - del_) * dz) * (qen - q[k]);
          tp += fe_ * (te - tp);
          qp += fe_ * (qen - qp);
          gk[k] += Mp * expm1((eps_ - del_) * dz);
          Mp *= exp((eps_ - del_) * dz);
        } else {
          const double g_ = expm1(par[P_ENT_CLOUD] * dz), fe = g_ / (1 + g_);
          if (par[P_ENT_DET] > 0) qa[k] += Mp * fe * (qp - q[k]);
          tp += fe * (te - tp);
          qp += fe * (q[k] - qp);
          if (!(par[P_ENT_DET] > 0)) Mp *= 1 + g_;
        }
      }
      if (par[P_PLUME_ENT] > 0) {
        qa[k] += Mp * f * (qp - q[k]);
        tp += f * (te - tp);
        qp += f * (q[k] - qp);
      }
      const double tpv = tp * VF(qp), tev = te * VF(q[k]);
      sum += mu[k] / Mk * (tpv - tev);
      last = tpv > tev;
      {
        const double Tp = tp * ex;
        if (tpv > tev) cape += rd * (tpv - tev) * ex * grav * mu[k] / Mk / pl;
        if (q0c < 0 && Tp < Tf) {
          q0c = qp;
          p0c = pl;
        }
        if (q40 < 0 && Tp < Tf - 40) {
          q40 = qp;
          p40 = pl;
        }
      }
      if (last) {
        *A = sum;
        *kd = k;
        *thd = tp;
        *qd = qp;
        *G = Mp;
        *FZ = fz;
        for (int m = 0; m <= bot; m++) qx[m] = qa[m];
        if (Gl)
          for (int m = 0; m <= bot; m++) Gl[m] = gk[m];
        if (ptop) *ptop = pl;
      }
    }
    /* (the shallow plume stops in the cloud layer; with cu_dth in the cumulus layer or the cloud layer above it) */
    if (!last || k == 0 || (shal && !(cu && k == bot - 1))) break;
    pt -= grav * mu[k];
    double ext;
    LIFT(pt, ext);
    if (!(tp * VF(qp) > b * VF(q[k - 1]))) break;
  }
#undef LIFT
#undef VF
  if (q0c >= 0 && q40 < 0) {
    q40 = *qd;
    p40 = pt;
  }
  lt[0] = q0c >= 0 ? fmax(0.0, q0c - q40) : 0.0;
  lt[1] = q0c >= 0 ? p0c : 0.0;
  lt[2] = q0c >= 0 ? p40 : 0.0;
  lt[3] = cape;
}

/* cu_top (air.cu:668-680): the theta the bottom layer's air reaches lifted from its top by cu_dp (Pa), condensing on
 * the way, in eight equal steps (no higher than 1000 Pa). */
static double cu_top(int L, const double* mu, const double* th, const double* q, const double* par, double eps,
                     int mix) {
  const int bot = L - 1;
  double pb = 0;
  for (int k = 0; k < bot; k++) pb += par[P_GRAV] * mu[k];
  double thp = th[bot], qp = q[bot];
  for (int s = 1; s <= 8; s++) {
    const double p = fmax(pb - par[P_CU_DP] * s / 8, 1000.0);
    condense(&thp, &qp, par[P_CP] * fpow_(p / par[P_P0], par[P_KAPPA], mix), p, par, eps, mix);
  }
  return thp;
}

/* The cumulus layer's floating edges (air.cu:913-922, again before the crossing at 1403-1407): the cumulus layer's
 * top (interface L - 3) at cu_dth above theta*, or with cu_dp at cu_top, kept 0.5 K above theta* and below both the
 * layer's mean less 1 K and the edge above less 0.5 K; with cl_float the cloud layer's lower edge (interface L - 4)
 * cl_float above cu_top where that is above its fixed value ths_cl, below the layer's mean less 1 K and the edge
 * above less 0.5 K. */
static void cu_edges(int L, const double* mu, const double* th, const double* q, double* par, double ths_cl,
                     double eps, int mix) {
  if (par[P_CL_FLOAT] > 0 && par[P_CU_DTH] > 0 && par[P_CU_DP] > 0 && L > 4)
    par[P_THS + L - 4] = fmax(ths_cl, fmin(cu_top(L, mu, th, q, par, eps, mix) + par[P_CL_FLOAT],
                                           fmin(th[L - 4] - 1, par[P_THS + L - 5] - 0.5)));
  if (par[P_CU_DTH] > 0 && L > 3) {
    const double top = par[P_CU_DP] > 0 ? cu_top(L, mu, th, q, par, eps, mix) : par[P_THETA_STAR] + par[P_CU_DTH];
    par[P_THS + L - 3] = fmax(par[P_THETA_STAR] + 0.5, fmin(top, fmin(th[L - 3] - 1, par[P_THS + L - 4] - 0.5)));
  }
}

/* plume_out (air.cu:682-713): a plume of base mass *m from the bottom layer (theta th0, vapor q0) detraining into
 * layer kd at (thp, qc), having taken in Gl[k] per kg from each layer k and given qx[k] by mixing. *m is first cut
 * to the layers' air above the floor and half their vapor. The air taken in below kd rises with it; kd's own leaves
 * at the layer's vapor and comes back at the plume's. cond: its condensate per kg (det of it left as cloud, at most
 * ldet per kg of plume); G its mass at kd per kg; the column's energy kept, the residual into kd. */
static void plume_out(int L, const double* par, double* mu, double* th, double* q, double* pib, double* cl, int kd,
                      double* m, double th0, double q0, double thp, double qc, const double* Gl, const double* qx,
                      double* wxo, double ldet, double* cond, double* det, double* G) {
  const int bot = L - 1;
  const double floor_ = par[P_FLOOR];
  double mm = fmin(*m, mu[bot] - floor_);
  for (int k = 0; k < bot; k++) {
    if (Gl[k] > 0) mm = fmin(mm, fmax(0.0, mu[k] - floor_) / Gl[k]);
    const double take = Gl[k] * q[k] - qx[k];
    if (take > 0) mm = fmin(mm, 0.5 * mu[k] * q[k] / take);
  }
  *G = 1;
  *cond = 0;
  *det = 0;
  if (!(mm > 0)) {
    *m = 0;
    return;
  }
  *m = mm;
  double qin = 0, xs = 0, Gt = 1;
  for (int k = 0; k < bot; k++) {
    Gt += Gl[k];
    qin += Gl[k] * q[k];
    xs += qx[k];
  }
  *G = Gt;
  const double cond_ = q0 + qin - Gt * qc - xs;
  *cond = fmax(0.0, cond_);
  *det = cl ? fmin(*cond, ldet * Gt) : 0.0;
  const double qarr = qc + (cond_ - *cond) / Gt;
  double tk[OG_LMAX], qk[OG_LMAX];
  for (int k = 0; k < bot; k++) {
    tk[k] = th[k];
    qk[k] = q[k];
  }
  const double Ec = column_energy(L, mu, pib, th, q, par);
  for (int k = 0; k < bot; k++) q[k] = fmax(0.0, q[k] + mm * qx[k] / mu[k]);
  move_mass(L, par, mu, th, q, pib, bot, kd, mm, th0, thp, q0, qarr, cl);
  for (int j = kd; j < bot; j++) wxo[j] += mm;
  for (int k = kd + 1; k < bot; k++)
    if (Gl[k] > 0) {
      move_mass(L, par, mu, th, q, pib, k, kd, mm * Gl[k], tk[k], thp, qk[k], qarr, cl);
      for (int j = kd; j < k; j++) wxo[j] += mm * Gl[k];
    }
  if (Gl[kd] > 0) q[kd] = fmax(0.0, q[kd] + mm * Gl[kd] * (qarr - qk[kd]) / mu[kd]);
  th[kd] += (Ec - column_energy(L, mu, pib, th, q, par)) / (mu[kd] * pib[kd]);
}

/* plume_ascent on its own, for the tests' closed-form checks of its options: out = (A, kd, thd, qd, G, FZ, lt[0..3],
 * ptop), qx (L) what its mixing gave each layer per kg, Gl (L) what it took in from each. */
void og_plume_ascent(int L, const double* par, const og_air* air, const double* mu, const double* th, const double* q,
                     double th0, double q0, double pi_, int shal, int mix, double* out, double* qx, double* Gl) {
  double A = 0, thd = 0, qd = 0, G = 1, FZ = 0, lt[4] = {0, 0, 0, 0}, ptop = pi_;
  int kd = -1;
  plume_ascent(L, mu, th, th0, q0, pi_, par, air, mix, &A, &kd, &thd, &qd, q, qx, lt, &G, &FZ, shal, Gl, &ptop);
  out[10] = ptop;
  out[0] = A;
  out[1] = kd;
  out[2] = thd;
  out[3] = qd;
  out[4] = G;
  out[5] = FZ;
  for (int m = 0; m < 4; m++) out[6 + m] = lt[m];
}

/* storm_flashes (air.cu:718-729): lightning from one storm's physics (Deierling et al. 2008): 9.0e-15 f_p f_np
 * flashes per minute, f_p the falling graupel's mass flux and f_np the rising small ice's, the plume's ice flux
 * shared among storms of updraft area storm_area rising at half sqrt(2 CAPE); only an updraft faster than graupel
 * falls holds it up to charge. Flashes per second per m^2. */
static double storm_flashes(double mflux, const double* lt, const double* par, const og_air* air, int mix) {
  const double qi = lt[0], w = 0.5 * sqrt(2 * fmax(0.0, lt[3]));
  const double gv = par[P_GRAUPEL_V], Tf = par[P_TF], rd = air->rd;
  if (!(mflux > 0) || !(qi > 0) || !(w > gv)) return 0;
  const double rho = lt[1] / (rd * Tf), phi = rho * w * par[P_STORM_AREA], lift = 1 - gv * gv / (w * w);
  const double lg = mix ? (double)logf((float)(lt[1] / lt[2])) : log(lt[1] / lt[2]);
  const double H = rd * (Tf - 20) / par[P_GRAV] * lg;
  const double up = phi * qi * lift, fnp = up * par[P_ICE_SMALL], fp = up * (1 - par[P_ICE_SMALL]) * H / w * gv;
  return mflux / phi * 9.0e-15 * fp * fnp / 60;
}

/* wet_bulb (air.cu:453-466): air (th, q) at pressure p evaporating water until saturated, three Newton steps with
 * Tetens' slope. */
static void wet_bulb(double* th, double* q, double p, const double* par, double eps, int mix) {
  const double ex = fpow_(p / par[P_P0], par[P_KAPPA], mix);
  for (int it = 0; it < 3; it++) {
    const double T = *th * ex;
    if (T < 160) return;
    const double qs = qsat_x(T, p, eps, par[P_ICE_AIR], mix), d = qs - *q;
    if (!(d > 0)) return;
    const double dqs = qs * 17.27 * (273.15 - 35.85) / ((T - 35.85) * (T - 35.85));
    const double e = d / (1 + par[P_LV] / par[P_CP] * dqs);
    *q += e;
    *th -= par[P_LV] * e / (par[P_CP] * ex);
  }
}

void og_air_column(int N, int L, const double* par_in, const og_air* air, int mode, const double* t,
                   const double* insol, const double* pos, const double* uabs, const double* mu_in, double* theta,
                   double* qv, double* Ts, double* ice, double dt, double* wx, double* olr, double* sw_abs,
                   double* stored, double* rain, double* evap, double* conv, double* tsurf, double* rainp,
                   double* cap, double* qocean, double* qsw, const double* elev, double* flash,
                   const double* radf, const double* swf, double* swalb, const double* ch, double* firn,
                   double* cloudw, double* cloudo, const double* land, const double* soilcap, double* soil,
                   double* pond, double* snow, const double* lc, double* xfer, double* dgo, int mix) {
#pragma omp parallel for schedule(static)
  for (int i = 0; i < N; i++) {
    double par[132 + OG_LMAX];
    memcpy(par, par_in, sizeof(par));
    double xfl[OG_GASN * OG_LMAX];                    /* the column's gas tracers, carried by every move of its air */
    for (int g = 0; g < OG_GASN; g++)
      for (int k = 0; k < L; k++) xfl[g * OG_LMAX + k] = xfer ? xfer[((size_t)g * L + k) * N + i] : 0.0;
    xf_col = xfer ? xfl : NULL;
    if (elev && par[P_THETA_RISE] > 0 && elev[i] > 0) par[P_THETA_STAR] += par[P_THETA_RISE] * elev[i];
    const double ts_base = par[P_THETA_STAR];
    par[P_BL_CAP] = omexp(dt / 86400.0, mix);       /* (the mixed layer takes in at most a day's share per step) */
    const int blm = par[P_BL_MIX] > 0;
    const double grav = par[P_GRAV], cp = par[P_CP], kappa = par[P_KAPPA], p0 = par[P_P0], sigma = par[P_SIGMA];
    const double eps_s = par[P_EPS_S], Tf = par[P_TF], Lf = par[P_LF], Lv = par[P_LV], aeps = air->eps;
    const int moist = par[P_MOIST] > 0;
    const int cloudy = moist && par[P_CLOUD] > 0 && cloudw;
    double mu[OG_LMAX], th[OG_LMAX], q[OG_LMAX], pb[OG_LMAX], pib[OG_LMAX], Q[OG_LMAX], T[OG_LMAX], eps[OG_LMAX];
    double cl[OG_LMAX], fc[OG_LMAX], liq[OG_LMAX], cwp[OG_LMAX], eps0[OG_LMAX];
    double mt = 0;
    for (int k = 0; k < L; k++) {
      mu[k] = mu_in[(size_t)k * N + i];
      th[k] = theta[(size_t)k * N + i];
      q[k] = moist ? qv[(size_t)k * N + i] : 0.0;
      mt += mu[k];
      Q[k] = 0;
      cl[k] = cloudy ? cloudw[(size_t)k * N + i] : 0.0;
      fc[k] = liq[k] = cwp[k] = 0;
    }
    layer_exner(L, mu, par, pb, pib);
    const int top = 0, bot = L - 1;
    const double exb = pow(pb[bot] / p0, kappa);     /* (p_bot / p0)^kappa at the column's lower edge */
    /* the mixed bottom layer: the lowest interface's theta* no colder than the layer, below the air above's mean */
    if (blm && L > 2) par[P_THETA_STAR] = fmin(fmax(ts_base, th[L - 1]), th[L - 2] - 1);
    const double ths_cl = L > 4 ? par[P_THS + L - 4] : 0.0;   /* (cl_float: the cloud layer's fixed lower edge) */
    cu_edges(L, mu, th, q, par, ths_cl, aeps, mix);
    double m_ice = moist ? ice[i] : 0.0;
    const int isl = moist && land && land[i] > 0.5;  /* land: the soil bucket, ponds, snow and the ground's heat */
    if (cap && par[P_C_WET] > 0 && isl) {            /* (c_wet: the ground's capacity by its water, its heat kept) */
      const double r = fmin(1.0, soil[i] / soilcap[i]), Cn = par[P_C_DRY] + (par[P_C_WET] - par[P_C_DRY]) * r;
      Ts[i] = Tf + (Ts[i] - Tf) * cap[i] / Cn;
      cap[i] = Cn;
    }
    const double Cw = cap ? cap[i] : par[P_C_SLAB];
    const double Cf = firn && par[P_FIRN] > 0 ? par[P_C_ICE] * par[P_RHO_FIRN] * par[P_FIRN] : 0.0;
    double Tfn = Cf > 0 ? firn[i] : Tf;
    const double E0 = column_energy(L, mu, pib, th, q, par) + Cw * Ts[i] - Lf * m_ice + Cf * (Tfn - Tf);
    for (int k = 0; k < L; k++) T[k] = pib[k] * th[k] / cp;
    /* each layer's cover, in-cloud water path (kg/m^2) and liquid share (all ice below -liq_cold C); a middle layer's
     * clouds sit high in it, so their phase is by its upper quarter's pressure */
    if (cloudy)
      for (int k = 0; k < L; k++) {
        const double ptk = k ? pb[k - 1] : 0.0;
        const double pc = k == bot ? 0.5 * (pb[k] + ptk) : ptk + 0.25 * (pb[k] - ptk);
        liq[k] = fmax(0.0, fmin(1.0, (th[k] * fpow_(pc / p0, kappa, mix) - Tf + par[P_LIQ_COLD]) / par[P_LIQ_COLD]));
        double qs = 0, dqs, l, f;
        if (!mixed_cloud(L, k, mu, th[k], q[k] + cl[k], par, aeps, mix, &l, &f, &dqs, 0)) {
          layer_cap(L, k, mu, th[k], par, aeps, mix, &qs, &dqs);
          cloud_pdf(q[k] + cl[k], qs, par, &l, &f);
        } else if (par[P_CW_QS] > 0) {
          double d_;
          layer_cap(L, k, mu, th[k], par, aeps, mix, &qs, &d_);
        }
        fc[k] = fmax(f, cover_of(cl[k], mu[k], liq[k], par, qs));
        cwp[k] = fc[k] > 1e-6 ? mu[k] * cl[k] / fc[k] : 0.0;
      }
    /* cu_cloud (with ml_clear and bl_q): the mixed layer clear below its cloud base, the cloud its thermals make is
     * above it, in the layer above: the thermals' statistics at the mixed layer's levels give its cover there and its
     * water (cu_cloud 2: the cover alone), for the radiation */
    if (cloudy && par[P_CU_CLOUD] > 0 && par[P_ML_CLEAR] > 0 && par[P_BL_Q] > 0 && L > 2) {
      double l_, f_, d_;
      if (mixed_cloud(L, bot, mu, th[bot], q[bot] + cl[bot], par, aeps, mix, &l_, &f_, &d_, 1) && f_ > 1e-6) {
        const int kc = bot - 1;
        const double w_ = mu[kc] * cl[kc] + (par[P_CU_CLOUD] > 1.5 ? 0.0 : mu[bot] * l_);
        fc[kc] = fmax(fc[kc], f_);
        cwp[kc] = w_ / fc[kc];
      }
    }
    /* the air touching the ground: the bottom mu_bl of the column (or bl_bot's share of it, the thin bottom layer's
     * own mass at least); the layers above give their air at their lower edge's theta* where warmer */
    double wbl[OG_LMAX], thb = 0, qb = 0, mb = 0;
    {
      const double mu_bl = par[P_MU_BL], bl_bot = par[P_BL_BOT];
      double need = blm ? mu[bot] : bl_bot > 0 && mu[bot] < mu_bl ? fmax(mu[bot], bl_bot * mu_bl) : mu_bl;
      for (int k = bot; k >= 0; k--) {
        const double m = mu[k] < need ? mu[k] : need;
        wbl[k] = m;
        need -= m;
        thb += m * (bl_bot > 0 && k < bot && th[k] > tstar(k, L, par) ? tstar(k, L, par) : th[k]);
        qb += m * q[k];
        mb += m;
      }
    }
    thb /= mb;
    qb /= mb;
    const double ps = pb[bot];
    const double Tsa = thb * exb;
    const int sfc = par[P_SFC] > 0 && ch;
    const double rough = !sfc && isl ? par[P_LAND_EXCH] : 1.0;
    const double bulk_c = par[P_BULK_C], gust = par[P_BULK_GUST], rd = air->rd;
    double stab = 1;
    if (bulk_c > 0 && par[P_BULK_Z] > 0) {
      const double Ub = uabs[i] + gust, Ri = grav * par[P_BULK_Z] * (Tsa - Ts[i]) / (Tsa * Ub * Ub);
      stab = Ri > 0 ? 1 / (1 + 10 * Ri * (1 + 5 * Ri)) : 1 - 10 * Ri / (1 + 10 * sqrt(-Ri));
    }
    double bulk = sfc ? ch[i] : bulk_c > 0 ? ps / (rd * Tsa) * bulk_c * stab * (uabs[i] + gust) : 0.0;
    if (!sfc && bulk_c > 0 && par[P_BULK_W] > 0) {
      const double gt = ps / (rd * Tsa) * par[P_BULK_W] * (Ts[i] > Tsa ? 3.0 : 1.0);
      bulk = bulk * gt / (bulk + gt);
    }
    if (sfc) bulk = bulk / (1 + bulk * dt / mb);
    const double hs = rough * (bulk_c > 0 || sfc ? cp * bulk : par[P_H_SENS] * (1 + uabs[i] / par[P_U0]));

    /* shortwave; the surface's albedo with the sea ice's cover */
    double qrad[OG_LMAX], qet[OG_LMAX], qeb[OG_LMAX];
    for (int k = 0; k < L; k++) qrad[k] = qet[k] = qeb[k] = 0;
    double I;
    if (mode == 0) {
      double fd;
      const double ang = og_sun_lon(t[0], par, &fd);
      const double dec = asin(sin(par[P_TILT]) * sin(ang));
      const double rot = 2 * 3.14159265358979323846 * fmod(t[0], par[P_DAY]) / par[P_DAY];
      const double sx = cos(dec) * cos(-rot), sy = cos(dec) * sin(-rot), sz = sin(dec);
      const double cz = pos[3 * i] * sx + pos[3 * i + 1] * sy + pos[3 * i + 2] * sz;
      I = par[P_S0] * fd * (cz > 0 ? cz : 0.0);
    } else {
      I = insol[i];
    }
    const double dice = m_ice / par[P_RHO_ICE];
    const double fice = moist ? fmin(1.0, dice / par[P_ICE_COVER]) : 0.0;
    /* the ground's albedo by its wetness (and the cover's lc rows: 1 dry, 2 wet, 3 the vegetated share, 4 the
     * vegetation's), its ponds at the sea's; the ice's, on land by its snow; lc row 5 the share of ice that shows */
    const double ice0 = m_ice;
    double wet = isl ? soil[i] : 0.0, pw = isl ? pond[i] : 0.0;
    const double pcov = isl && pw > 0 ? fmin(1.0, pw / par[P_POND_COVER]) : 0.0;
    const size_t Ns = (size_t)N;
    const double alb0 =
        !isl ? par[P_ALB]
        : lc ? pcov * par[P_ALB] + (1 - pcov) * (lc[3 * Ns + i] * (lc[Ns + i] + (lc[2 * Ns + i] - lc[Ns + i]) * wet /
                                                                               soilcap[i]) +
                                                 (1 - lc[3 * Ns + i]) * lc[4 * Ns + i])
             : pcov * par[P_ALB] + (1 - pcov) * (par[P_ALB_DRY] + (par[P_ALB_WET] - par[P_ALB_DRY]) * wet / soilcap[i]);
    const double warm = fmax(0.0, fmin(1.0, (Tsa - Tf + par[P_ALB_RAMP]) / par[P_ALB_RAMP]));
    double alb_i = par[P_ALB_ICE] + (par[P_ALB_ICE_MELT] - par[P_ALB_ICE]) * warm;
    if (isl) {
      const double fs = fmin(1.0, snow[i] / par[P_SNOW_COVER]);
      alb_i = fs * (par[P_ALB_SNOW] + (par[P_ALB_SNOW_MELT] - par[P_ALB_SNOW]) * warm) + (1 - fs) * par[P_ALB_GLACIER];
    }
    const double fsee = isl && lc ? fice * lc[5 * Ns + i] : fice;
    const double alb = fsee * alb_i + (1 - fsee) * alb0;
    if (swalb) swalb[i] = alb;
    double toAir;
    if (swf) {
      double s = 0;
      for (int k = 0; k < L; k++) s += swf[(size_t)k * N + i];
      toAir = s * I;
    } else {
      toAir = par[P_ASW] * (mt / par[P_MU_REF]) * I;
    }
    if (swf) {
      for (int k = 0; k < L; k++) {
        Q[k] += I * swf[(size_t)k * N + i];
        qrad[k] += I * swf[(size_t)k * N + i];
        qet[k] += I * swf[(size_t)(L + 2 + 2 * k) * N + i];
        qeb[k] += I * swf[(size_t)(L + 3 + 2 * k) * N + i];
      }
    } else if (par[P_SW_TAU] > 1) {
      double st = 0;
      for (int k = 0; k < L; k++) {
        const double sb = pb[k] / pb[bot], sb4 = sb * sb * sb * sb;
        Q[k] += toAir * (sb4 - st);
        st = sb4;
      }
    } else if (par[P_SW_TAU] > 0) {
      for (int k = 0; k < L; k++)
        Q[k] += toAir * (tau_p(pb[k], pb[bot], par) - tau_p(k ? pb[k - 1] : 0.0, pb[bot], par)) / par[P_TAU0];
    } else {
      for (int k = 0; k < L; k++) Q[k] += toAir * mu[k] / mt;
    }
    /* clouds reflect: each layer's by conservative two-stream scattering of its optical depth 3 W / (2 rho r), the
     * layers at random overlap, with the reflections between them and the ground */
    double Tc = 1;
    if (cloudy)
      for (int k = 0; k < L; k++) {
        const double tau = 1.5 * cwp[k] * (liq[k] / (1000 * par[P_RE_LIQ]) + (1 - liq[k]) / (par[P_RHO_ICE] * par[P_RE_ICE]));
        const double g1 = (1 - par[P_CLOUD_G]) * tau;
        Tc *= 1 - fc[k] * g1 / (2 + g1);
      }
    const double swS = swf ? I * swf[(size_t)L * N + i] : (I - toAir) * Tc * (1 - alb) / (1 - alb * (1 - Tc));

    /* longwave: each layer's emissivity from its share of the optical depth (moist with lw_w0: the vapor's part from
     * the column's own vapor, growing as the square root of its path) */
    if (moist && par[P_LW_W0] > 0) {
      const double tau0 = par[P_TAU0], tau_lin = par[P_TAU_LIN];
      double W = 0;
      for (int k = 0; k < L; k++) W += mu[k] * (q[k] > 0 ? q[k] : 0.0);
      const double tv = tau0 * (1 - tau_lin) * sqrt(W / par[P_LW_W0]);
      for (int k = 0; k < L; k++) {
        const double sb = pb[k] / pb[bot], st = k ? pb[k - 1] / pb[bot] : 0.0;
        eps[k] = omexp((tau0 * tau_lin * (sb - st) + (W > 0 ? tv * mu[k] * (q[k] > 0 ? q[k] : 0.0) / W : 0.0)), mix);
      }
    } else {
      for (int k = 0; k < L; k++)
        eps[k] = omexp(tau_p(pb[k], pb[bot], par) - tau_p(k ? pb[k - 1] : 0.0, pb[bot], par), mix);
    }
    if (cloudy)                                      /* clouds absorb and emit by their water path */
      for (int k = 0; k < L; k++) {
        eps0[k] = eps[k];
        const double kl = liq[k] * par[P_K_LW_LIQ] + (1 - liq[k]) * par[P_K_LW_ICE];
        eps[k] = 1 - (1 - eps[k]) * (1 - fc[k] * omexp(kl * cwp[k], mix));
      }
    double ldown = 0;
    const int RW = L + 1;
    if (radf) {
      ldown = radf[(size_t)(2 * RW - 1) * N + i] - radf[(size_t)L * N + i];
    } else {
      for (int k = 0; k < L; k++) {
        double dn = eps[k] * sigma * T[k] * T[k] * T[k] * T[k];
        for (int j = k + 1; j < L; j++) dn -= eps[j] * dn;
        ldown += dn;
      }
    }
    /* the surface: open water at Ts, or under ice thicker than ice_cover a skin balancing conduction through the ice
     * against sunlight, longwave and sensible heat, capped at freezing */
    const double kex = rough * (bulk_c > 0 || sfc ? bulk : par[P_K_EVAP] * (1 + uabs[i] / par[P_U0]));
    /* q_shape: the air at the surface holds the bottom layer's vapor at its relative humidity there (its capacity
     * layer_cap's), not the layer's mean; bl_q keeps the layer's vapor even */
    double qa = qb;
    if (moist && par[P_Q_SHAPE] > 0 && !(par[P_BL_Q] > 0)) {
      double qsl, dqsl;
      layer_cap(L, bot, mu, th[bot], par, aeps, mix, &qsl, &dqsl);
      const double qss = qsat_x(Tsa, ps, aeps, 0.0, mix);
      if (qsl > 0) qa = fmin(qb * qss / qsl, qss);
    }
    /* soil and plants hold their water back: a surface resistance rs in series with the air's ra (land_rs; rs_max:
     * the stomata open with the sunlight, rs = land_rs (1 + f) / (land_rs / rs_max + f), f the sunlight reaching the
     * ground over rs_sun; lc rows 6, 7 the cover's own land_rs, rs_sun; its bare share, row 3's complement, has
     * none); bare_fx: the bare soil's evaporation falls as (wet / full)^bare_fx, the plants' as wet / full */
    double rfac = 1, fbare = 0, rveg = 1;
    if (isl && par[P_LAND_RS] > 0 && kex > 0) {
      const double r0 = lc ? lc[6 * Ns + i] : par[P_LAND_RS], rsun = lc ? lc[7 * Ns + i] : par[P_RS_SUN];
      double rs = r0;
      if (par[P_RS_MAX] > 0) {
        const double f = swS / fmax(1 - alb, 0.05) / rsun;
        rs = r0 * (1 + f) / (r0 / par[P_RS_MAX] + f);
      }
      const double ra = ps / (rd * Tsa) / kex;
      rfac = ra / (ra + rs);
      rveg = rfac;
      if (lc) {
        fbare = lc[3 * Ns + i];
        rfac = fbare + (1 - fbare) * rfac;
      }
    }
    const int bfx = isl && par[P_BARE_FX] > 0 && lc && par[P_LAND_RS] > 0;
#define SOILF(w_) (fbare * pow(fmin(1.0, (w_)), par[P_BARE_FX]) + (1 - fbare) * rveg * fmin(1.0, (w_)))
    const double qfac = !isl && par[P_Q_SEA] > 0 ? par[P_Q_SEA] : 1.0;
    const int skin = moist && dice > par[P_ICE_COVER];
    /* firn where the land's ice is thick enough; where it ends, its heat goes to the ground (or the slab) below */
    const int fon = Cf > 0 && isl && skin && m_ice >= par[P_RHO_FIRN] * par[P_FIRN];
    if (Cf > 0 && !fon && Tfn != Tf) {
      Ts[i] += Cf * (Tfn - Tf) / Cw;
      Tfn = Tf;
    }
    const double ts = Ts[i];
    double tsk = ts, melt_top = 0, cond = 0, condb = 0;
    if (skin) {
      /* the skin conducts to what lies under it: the ice down to the ground (or the slab), or the firn's middle */
      const double kc = fon ? 2 * par[P_K_FIRN] / par[P_FIRN] : par[P_K_ICE] / dice, tb = fon ? Tfn : ts;
      double x = fmin(tb, Tf);
      for (int it = 0; it < 8; it++) {
        const double x3 = x * x * x;
        const double f = kc * (tb - x) - eps_s * sigma * x3 * x + swS + ldown - hs * (x - Tsa);
        x -= f / (-kc - 4 * eps_s * sigma * x3 - hs);
      }
      if (x > Tf) x = Tf;
      const double x4 = x * x * x * x;
      melt_top = kc * (tb - x) - eps_s * sigma * x4 + swS + ldown - hs * (x - Tsa);
      tsk = x;
      cond = kc * (tb - tsk);
      if (fon) {                                     /* and from the ground through the rest of the ice to the firn */
        const double rest = fmax(dice - par[P_RHO_FIRN] * par[P_FIRN] / par[P_RHO_ICE], 0.0);
        condb = (ts - Tfn) / (rest / par[P_K_ICE] + par[P_FIRN] / (2 * par[P_K_FIRN]));
      }
    }
    /* the ground's skin (c_deep; its capacity Cw): stepped implicitly, C (x - ts) = dt F(x) with F the surface's net
     * gain (sunlight, longwave, sensible heat, the evaporation before the water runs short), by Newton's method */
    if (isl && !skin && par[P_C_DEEP] > 0) {
      const double wfull = wet / fmax(par[P_SOIL_BETA] * soilcap[i], 1e-9);
      const double beta = kex * (pcov + (1 - pcov) * (bfx ? SOILF(wfull) : rfac * fmin(1.0, wfull))) * (1 - fice);
      double x = ts;
      for (int it = 0; it < 6; it++) {
        const double x3 = x * x * x, d = qsat_x(x, ps, aeps, 0.0, mix) - qa;
        const double dqs = mix ? (double)((float)qsat_x(x + 0.01, ps, aeps, 0.0, 1) - (float)qsat_x(x - 0.01, ps, aeps, 0.0, 1))
                               : qsat_x(x + 0.01, ps, aeps, 0.0, 0) - qsat_x(x - 0.01, ps, aeps, 0.0, 0);
        const double e = d > 0 ? beta * d : 0.0, de = d > 0 ? beta * dqs * 50 : 0.0;
        const double g = Cw * (x - ts) - dt * (swS + ldown - eps_s * sigma * x3 * x - hs * (x - Tsa) - Lv * e);
        x -= g / (Cw + dt * (4 * eps_s * sigma * x3 + hs + Lv * de));
      }
      tsk = x;
    }
    double up = eps_s * sigma * tsk * tsk * tsk * tsk;
    if (radf) {
      const double Es = up;
#define RF(r, k) radf[(size_t)((r) * RW + (k)) * N + i]
      for (int k = 0; k < L; k++) {
        const double ql = (RF(0, k + 1) - RF(1, k + 1)) - (RF(0, k) - RF(1, k)) + (RF(2, k + 1) - RF(2, k)) * Es;
        Q[k] += ql;
#define RE(r) radf[(size_t)(3 * RW + 3 + 6 * k + (r)) * N + i]
        const double n1 = RE(0) - RE(1) + RE(2) * Es, n2 = RE(3) - RE(4) + RE(5) * Es;
#undef RE
        qrad[k] += ql;
        qet[k] += n1 - ((RF(0, k) - RF(1, k)) + RF(2, k) * Es);
        qeb[k] += ((RF(0, k + 1) - RF(1, k + 1)) + RF(2, k + 1) * Es) - n2;
      }
      up = RF(0, 0) + RF(2, 0) * Es;
      if (cloudy && cloudo) {                        /* cover, the clouds' effects at the top, optical depths */
        double clear = 1;
        for (int k = 0; k < L; k++) {
          cloudo[(size_t)k * N + i] = fc[k];
          clear *= 1 - fc[k];
        }
        cloudo[(size_t)L * N + i] = 1 - clear;
        cloudo[(size_t)(L + 1) * N + i] =
            swf ? I * (swf[(size_t)(L + 1) * N + i] - (1 - (toAir / fmax(I, 1e-30) + swf[(size_t)L * N + i])))
                : swS - (I - toAir) * (1 - alb);
        cloudo[(size_t)(L + 2) * N + i] = radf[(size_t)(3 * RW) * N + i] + radf[(size_t)(3 * RW + 1) * N + i] * Es - up;
        for (int k = 0; k < L; k++) {
          const double re_l = 1000 * par[P_RE_LIQ], re_i = par[P_RHO_ICE] * par[P_RE_ICE];
          cloudo[(size_t)(L + 3 + k) * N + i] =
              fc[k] > 1e-6 ? (liq[k] * par[P_K_LW_LIQ] + (1 - liq[k]) * par[P_K_LW_ICE]) * cwp[k] : 0.0;
          cloudo[(size_t)(2 * L + 3 + k) * N + i] = 1.5 * cwp[k] * (liq[k] / re_l + (1 - liq[k]) / re_i);
          const double tl_ = liq[k] / re_l, ti_ = (1 - liq[k]) / re_i;
          cloudo[(size_t)(3 * L + 3 + k) * N + i] = tl_ + ti_ > 0 ? tl_ / (tl_ + ti_) : 1.0;
        }
      }
#undef RF
    } else {
      for (int k = bot; k >= 0; k--) {
        const double a = eps[k] * up;
        Q[k] += a;
        const double em = eps[k] * sigma * T[k] * T[k] * T[k] * T[k];
        Q[k] -= 2 * em;
        up = up - a + em;
      }
      if (cloudy && cloudo) {                        /* cover and the clouds' effects at the top */
        double up0 = eps_s * sigma * tsk * tsk * tsk * tsk, clear = 1;
        for (int k = bot; k >= 0; k--) up0 = up0 - eps0[k] * up0 + eps0[k] * sigma * T[k] * T[k] * T[k] * T[k];
        for (int k = 0; k < L; k++) {
          cloudo[(size_t)k * N + i] = fc[k];
          clear *= 1 - fc[k];
        }
        cloudo[(size_t)L * N + i] = 1 - clear;
        cloudo[(size_t)(L + 1) * N + i] = swS - (I - toAir) * (1 - alb);
        cloudo[(size_t)(L + 2) * N + i] = up0 - up;
      }
      for (int k = 0; k < L; k++) {
        double dn = eps[k] * sigma * T[k] * T[k] * T[k] * T[k];
        for (int j = k + 1; j < L; j++) {
          const double a = eps[j] * dn;
          Q[j] += a;
          dn -= a;
        }
      }
    }
    olr[i] = up;
    /* sensible heat and evaporation into the bottom air */
    const double H = hs * (tsk - Tsa);
    for (int k = 0; k < L; k++) Q[k] += H * wbl[k] / mb;
    double E = 0;
    if (moist) {
      const double def = qfac * qsat_x(tsk, ps, aeps, 0.0, mix) - qa;
      E = def > 0 ? kex * def * (1 - fice) : 0.0;
      if (isl) {                                     /* land: from its ponds freely, from the soil held back */
        const double full = par[P_SOIL_BETA] * soilcap[i];
        double Es = (1 - pcov) * E, Ep = pcov * E;
        if (bfx) {
          Es *= SOILF(wet / full);
        } else {
          Es *= rfac;
          if (wet < full) Es *= wet / full;
        }
        if (Es * dt > wet) Es = wet / dt;
        if (Ep * dt > pw) Ep = pw / dt;
        E = Es + Ep;
        wet -= Es * dt;
        pw -= Ep * dt;
      }
      for (int k = 0; k < L; k++) q[k] += E * dt * wbl[k] / mb / mu[k];
    }
    evap[i] = E;
    tsurf[i] = tsk;
    double dth[OG_LMAX], dlat[OG_LMAX];
    /* the diagnostics (dgo, air.cu:1286-1290): each layer's latent heating (W/m^2), the crossing's gross up and down
     * mass per interface, the convective mass delivered into each layer */
    double dgl[OG_LMAX], gross[2 * OG_LMAX], cin[OG_LMAX];
    for (int k = 0; k < L; k++) dgl[k] = cin[k] = gross[2 * k] = gross[2 * k + 1] = 0;
    for (int k = 0; k < L; k++) {
      dth[k] = Q[k] * dt / (mu[k] * pib[k]);
      th[k] += dth[k];
      dlat[k] = 0;
    }
    /* the slab: open water takes the surface's fluxes (evaporation costs it Lv per kg); under ice the conduction up
     * into the skin. Below Tf the slab's heat freezes into ice; ice melts with slab heat above Tf and from the top */
    double qS;
    if (fon) Tfn += (condb - cond) * dt / Cf;       /* the firn between the ground's conduction and the skin's */
    if (skin) qS = fon ? -condb : -cond;
    else qS = swS + ldown - eps_s * sigma * tsk * tsk * tsk * tsk - H - Lv * E;
    double Tn = ts + qS * dt / Cw;
    if (qocean) qocean[i] = qS;
    if (qsw) qsw[i] = skin ? 0.0 : swS;
    if (moist) {
      m_ice -= melt_top * dt / Lf;
      if (m_ice < 0) {
        Tn += -m_ice * Lf / Cw;
        m_ice = 0;
      }
      if (Tn < Tf) {
        if (!isl) {                                  /* (the ground cools below freezing; the sea freezes) */
          m_ice += (Tf - Tn) * Cw / Lf;
          Tn = Tf;
        }
      } else if (m_ice > 0) {
        const double m = (Tn - Tf) * Cw / Lf;
        if (m <= m_ice) {
          m_ice -= m;
          Tn = Tf;
        } else {
          Tn = Tf + (m - m_ice) * Lf / Cw;
          m_ice = 0;
        }
      }
      ice[i] = m_ice;
    }
    Ts[i] = Tn;
    sw_abs[i] = toAir + swS;
    /* condensation: one-step saturation adjustment per layer; its latent heat counts as heating for the crossing */
    double R = 0, Rk[2] = {0, 0}, Rc = 0, S = 0;
    const int icy = par[P_ICE_AIR] > 0;
    if (cloudy) {
      /* cloud water forms and erodes; liquid turns to rain (Sundqvist 1989), ice falls out; rain evaporates in the
       * clear air it falls through. With ice_air the falling ice freezes in its layer (Lf), melts in a layer whose
       * lower edge is above freezing, sublimates with the rain's evaporation; what is left reaches the ground as
       * snow. */
      double Fr = 0, Fs = 0, fa = 0, psf = 0;
      for (int k = 0; k < L; k++) psf += grav * mu[k];
      for (int k = 0; k < L; k++) {
        const double dl = cloud_adjust(L, k, mu, th, q, cl, pib, dt, par, aeps, mix);
        dth[k] += Lv * dl / pib[k];
        dlat[k] += Lv * dl / pib[k];
        dgl[k] += Lv * mu[k] * dl / dt;
        double qs, dqs, l, f;
        layer_cap(L, k, mu, th[k], par, aeps, mix, &qs, &dqs);
        double dqf;
        if (!mixed_cloud(L, k, mu, th[k], q[k] + cl[k], par, aeps, mix, &l, &f, &dqf, 0))
          cloud_pdf(q[k] + cl[k], qs, par, &l, &f);
        f = fmax(f, cover_of(cl[k], mu[k], liq[k], par, qs));
        const double inc = f > 1e-6 ? cl[k] / f * fmax(1.0, mu[k] / par[P_MU_CLOUD]) : 0.0, r = inc / par[P_Q_AUTO];
        double pt_ = 0;
        for (int j = 0; j < k; j++) pt_ += grav * mu[j];
        const double pm = pt_ + 0.5 * grav * mu[k], Tk = pib[k] * th[k] / cp, depth = mu[k] / (pm / (air->rd * Tk));
        const double rate = liq[k] * par[P_C_AUTO] * omexp(r * r, mix) + (1 - liq[k]) * par[P_V_ICE] / depth;
        const double loss = cl[k] * omexp(rate * dt, mix);
        cl[k] -= loss;
        Fr += mu[k] * loss / dt;
        if (icy) {
          const double fz = rate > 0 ? fmin(1.0, (1 - liq[k]) * par[P_V_ICE] / depth / rate) : 0.0, dz = loss * fz;
          Fs += mu[k] * dz / dt;
          th[k] += Lf * dz / pib[k];
          dth[k] += Lf * dz / pib[k];
          dlat[k] += Lf * dz / pib[k];
          dgl[k] += Lf * mu[k] * dz / dt;
          if (Fs > 0 && base_T(L, k, mu, th, par, mix) > Tf) {
            const double dm = Fs * dt / (mu[k] * pib[k]) * Lf;
            th[k] -= dm;
            dth[k] -= dm;
            dlat[k] -= dm;
            dgl[k] -= Lf * Fs;
            Fs = 0;
          }
        }
        double ev = 0;
        if (loss > 0) fa = fmax(fa, f);
        const double rc_ = par[P_REVAP] > 0 ? (isl ? par[P_REVAP_RHL] : par[P_REVAP_RHS]) : 1.0;
        if (Fr > 0 && rc_ * qs > q[k]) {
          if (par[P_REVAP] > 0) {
            const double a_ = fmax(fa, par[P_REVAP_AREA]), ac = fmax(0.0, a_ - f);
            ev = ac * 5.44e-4 * (rc_ * qs - q[k]) * pow(sqrt(pm / psf) / 5.09e-3 * Fr / a_, 0.577) * mu[k];
          } else {
            ev = par[P_K_REVAP] * (1 - f) * (1 - q[k] / qs) * sqrt(Fr) * mu[k];
          }
          ev = fmin(ev, fmin(Fr, (rc_ * qs - q[k]) * mu[k] / dt / (1 + Lv / pib[k] * dqs)));
          const double es = Fr > 0 ? ev * Fs / Fr : 0.0;
          const double Le = Lv * ev + Lf * es;
          Fr -= ev;
          Fs -= es;
          q[k] += ev * dt / mu[k];
          th[k] -= Le * dt / (mu[k] * pib[k]);
          dth[k] -= Le * dt / (mu[k] * pib[k]);
          dlat[k] -= Le * dt / (mu[k] * pib[k]);
          dgl[k] -= Le;
        }
        Rk[k == bot ? 1 : 0] += mu[k] * loss - ev * dt;
      }
      R = Fr * dt;
      S = fmax(0.0, fmin(Fs * dt, R));
    } else if (moist)
      for (int k = 0; k < L; k++) {
        const double dq = condense_k(L, k, mu, th, q, pib, par, aeps, mix);
        dth[k] += Lv * dq / pib[k];
        dlat[k] += Lv * dq / pib[k];
        dgl[k] += Lv * mu[k] * dq / dt;
        R += mu[k] * dq;
        Rk[k == bot ? 1 : 0] += mu[k] * dq;
      }
    /* the crossing, with the k-tables' edge sub-levels' own heating at the layers' edges */
    if (blm && L > 2) par[P_THETA_STAR] = fmin(fmax(ts_base, th[bot]), th[bot - 1] - 1);
    cu_edges(L, mu, th, q, par, ths_cl, aeps, mix);
    /* (bl_lcl: the mixed layer ends at its condensation level; lcl_mass: the layer's mass up to where its air, lifted
     * dry-adiabatically, saturates, by bisection in pressure) */
    double mu_lcl = 1e30, bl_room = 1e30;
    if (par[P_BL_LCL] > 0 && blm && L > 2) {
      mu_lcl = lcl_mass(L, mu, th, q, par, aeps, mix);
      bl_room = fmax(0.0, mu_lcl - mu[bot]);
    }
    double dtop[OG_LMAX], dbot[OG_LMAX], dtop1[OG_LMAX], wxo[OG_LMAX];
    if (radf) {
      double pt_ = 0;
      for (int k = 0; k < L; k++) {
        const double pbk = pt_ + grav * mu[k], dp = pbk - pt_;
        const double ft = k == 0 ? 0.05 : 0.25, fb = k == 0 ? 0.5 : 0.25;
        const double pit = cp * fpow_((pt_ + 0.5 * ft * dp) / p0, kappa, mix);
        const double pibt = cp * fpow_((pbk - 0.5 * fb * dp) / p0, kappa, mix);
        const double mean = qrad[k] * dt / (mu[k] * pib[k]);
        dtop[k] = dth[k] - mean + qet[k] * dt / (ft * mu[k] * pit);
        dbot[k] = dth[k] - mean + qeb[k] * dt / (fb * mu[k] * pibt);
        if (k == L - 1 && blm) dtop[k] = dbot[k] = dth[k];   /* (the mixed bottom layer heats evenly) */
        pt_ = pbk;
      }
      if (par[P_DRY_EDGE] > 0 && L > 2) dtop[(int)par[P_EDGE_K]] -= dlat[(int)par[P_EDGE_K]];
    }
    const int edges = par[P_EDGE_HEAT] > 0;
    if (radf && !edges && L > 1) {
      for (int k = 0; k < L; k++) dtop1[k] = dth[k];
      for (int k = 1; k <= par[P_EDGE_K] && k < L; k++) dtop1[k] = dtop[k];
    }
    crossing(L, par, mu, th, q, dth, dth[top], pib, wxo,
             radf ? (edges ? dtop : (L > 1 ? dtop1 : NULL)) : NULL, edges && radf ? dbot : NULL, cloudy ? cl : NULL,
             bl_room, aeps, mix, gross);
    /* the mixed boundary layer: the surface's buoyancy flux sets its depth. Rising it entrains the air above at bl_ent
     * of that flux across the jump to it (Tennekes 1973), a slice from the layer's bottom edge; falling, it gives its
     * top back down to a thin night layer mu_night. bl_reach: dry convection first takes in the air above colder
     * than it, layer by layer while they empty. bl_lcl: it ends at its condensation level, the air above that leaving
     * into the layer above. */
    if (blm && L > 2) {
      int ka = bot - 1;
#define REACH_EDGE(k_) ((k_) < bot - 1 ? fmax(tstar((k_), L, par), th[bot]) : tstar((k_), L, par))
      if (par[P_BL_REACH] > 0) {
        for (int k = bot - 1; k >= 1; k--) {
          const double avail = mu[k] - par[P_FLOOR];
          if (avail > 0) {
            const double e = REACH_EDGE(k), D = th[k] - e;
            const int spread = D > 0 && th[k] < tstar(k - 1, L, par);
            const double x = spread ? (th[bot] > e ? mu[bot] / mu[k] * (sqrt(1 + mu[k] * (th[bot] - e) / (mu[bot] * D)) - 1) : 0.0)
                                    : (th[k] < th[bot] ? 1.0 : 0.0);
            const double want = fmin(1.0, x) * mu[k];
            const double m = fmin(fmin(fmin(want, avail), par[P_BL_CAP] * fmax(mu[k] - par[P_MU_BL], par[P_MU_BL])),
                                  fmax(0.0, mu_lcl - mu[bot]));
            if (m > 0) {
              const double fm = m / mu[k], thl = spread ? e + fm * D : th[k];
              const double qx = par[P_EDGE_VAP] > 0 ? edge_q(L, k, 0, mu, th, q, m, par, aeps, mix) : q[k];
              move_mass(L, par, mu, th, q, pib, k, bot, m, thl, thl, qx, qx, cloudy ? cl : NULL);
              for (int j = k; j < bot; j++) wxo[j] -= m;
            }
            if (!(m >= avail * (1 - 1e-9))) break;
          }
        }
        ka = bot - 1;
        while (ka > 1 && !(mu[ka] > par[P_FLOOR] * 1.000001)) ka--;
      }
      const double eka = REACH_EDGE(ka);
#undef REACH_EDGE
      const double Bf = H + air->vf * cp * Tsa * E, jump = fmax(1.0, eka - th[bot]);
      if (Bf > 0) {
        const double m = fmin(fmin(par[P_BL_ENT] * Bf * dt / (cp * jump), mu[ka] - par[P_FLOOR]),
                              par[P_ENT_LCL] > 0 ? 1e30 : fmax(0.0, mu_lcl - mu[bot]));
        if (m > 0) {
          const int spread = ka >= 1 && th[ka] < tstar(ka - 1, L, par) && th[ka] > eka;
          const double fm = fmin(1.0, m / mu[ka]);
          const double thc = spread ? eka + fm * (th[ka] - eka) : th[ka];
          const double qx = par[P_EDGE_VAP] > 0 ? edge_q(L, ka, 0, mu, th, q, m, par, aeps, mix) : q[ka];
          move_mass(L, par, mu, th, q, pib, ka, bot, m, thc, thc, qx, qx, cloudy ? cl : NULL);
          for (int j = ka; j < bot; j++) wxo[j] -= m;
        }
      }
      if (!(Bf > 0) && mu[bot] > par[P_MU_NIGHT]) {
        const double m = fmin(-Bf * dt / (cp * jump), mu[bot] - par[P_MU_NIGHT]);
        const double qx = par[P_EDGE_VAP] > 0 ? edge_q(L, bot, 1, mu, th, q, m, par, aeps, mix) : q[bot];
        move_mass(L, par, mu, th, q, pib, bot, bot - 1, m, th[bot], th[bot], qx, qx, cloudy ? cl : NULL);
        wxo[bot - 1] += m;
      }
      if (par[P_BL_LCL] > 0) {
        const double m = mu[bot] - fmax(lcl_mass(L, mu, th, q, par, aeps, mix), par[P_MU_NIGHT]);
        if (m > 0) {
          const double qx = par[P_EDGE_VAP] > 0 ? edge_q(L, bot, 1, mu, th, q, m, par, aeps, mix) : q[bot];
          move_mass(L, par, mu, th, q, pib, bot, bot - 1, m, th[bot], th[bot], qx, qx, cloudy ? cl : NULL);
          wxo[bot - 1] += m;
        }
      }
    }

    /* moist convection: a parcel of bottom-layer air lifted to the interface above it; buoyant there (theta_p above
     * theta*), it rises through every interface it stays buoyant at and detrains above the last, its condensate
     * raining; the mass that would remove the buoyancy at the first interface relaxes over tau_conv */
    double mc = 0, Scg = 0, qss_r = 0, ex_r = -1, ws_r = 0, rho_r = 0;   /* (the root's saturation, the thermals'
                                                                           * excess, w* and density: the shallow plume's) */
    const double mu_pre = mu[bot];                    /* (cu_sub: the mixed layer's mass before the plumes) */
    if (flash) flash[i] = 0;
    if (moist && par[P_THETA_STAR] > 0) {
      double pi_ = 0;
      for (int k = 0; k < bot; k++) pi_ += grav * mu[k];
      const double th0 = th[bot];
      double q0 = q[bot];
      if (par[P_PLUME_ROOT] > 0) {
        /* plume_root: plumes root in the air at the surface, the layer's vapor at its relative humidity there (edge_q
         * at its bottom edge); root_q: plus the thermals' excess b E / (rho w_m), b = 8.5, w_m = (u*^3 + 0.28
         * w*^3)^(1/3) (Holtslag & Boville 1993), up to the surface air's saturation; else with clouds its moistest
         * part by the statistics, against the surface air's capacity */
        const double ps0 = pi_ + grav * mu[bot], Ts0 = th0 * fpow_(ps0 / p0, kappa, mix);
        const double qss = Ts0 < 160 ? 0.0 : qsat_x(Ts0, ps0, aeps, par[P_ICE_AIR], mix);
        q0 = edge_q(L, bot, 0, mu, th, q, 0, par, aeps, mix);
        if (par[P_ROOT_Q] > 0) {
          const double rho = ps / (rd * Tsa), Ub = uabs[i] + gust;
          const double us3 = sfc ? fpow_(fmax(0.0, ch[i] / rho * Ub), 1.5, mix) : 0.0;
          const double Bk = fmax(0.0, (H + air->vf * cp * Tsa * E) / (rho * cp));
          const double hb = rd * Tsa / grav * log(ps / fmax(ps - grav * mu[bot], 1.0));
          const double wm = cbrt(us3 + 0.28 * grav / Tsa * Bk * hb);
          const double ex = wm > 1e-3 ? 8.5 * fmax(0.0, E) / (rho * wm) : 0.0;
          q0 += fmin(ex, fmax(0.0, qss - q0));
          qss_r = qss;
          ex_r = ex;
          rho_r = rho;
          ws_r = cbrt(grav / Tsa * Bk * hb);
        } else if (cloudy) {
          q0 = fmax(q0, fmin(qss, q0 + (1 - par[P_RH_CRIT]) * qss));
        }
      } else if (cloudy) {                           /* plumes root in the cell's moist air: its moistest part */
        double qs, dqs;
        layer_cap(L, bot, mu, th0, par, aeps, mix, &qs, &dqs);
        q0 = fmax(q0, fmin(qs, q0 + (1 - par[P_RH_CRIT]) * qs));
      }
      double thp, qc, thp2, qc2;
      if (par[P_PLUME_DEEP] > 0 && bot >= 2) {
        /* the deep plume: its full ascent (plume_ascent) against the layers' spread air, mixing and entraining; the
         * air it rises through sinks to make room, warming each level by the theta it had below, so the mass that
         * removes its buoyancy is A / (theta_d - theta*), relaxed over tau_conv; it takes in the cloud layer's air it
         * entrained (me), keeps l_det of its condensate as cloud, rains the rest (snow by its frozen share), and its
         * rain feeds a downdraft of the cloud layer's air, saturated by evaporating it */
        double A, G = 1, FZ = 0, qx[OG_LMAX], lt[4], xs = 0, Gl[OG_LMAX];
        int kd;
        plume_ascent(L, mu, th, th0, q0, pi_, par, air, mix, &A, &kd, &thp, &qc, q, qx, lt, &G, &FZ, 0, Gl, NULL);
        for (int k = 0; k <= bot; k++) xs += qx[k];
        if (kd >= 0 && A > 0) {
          const double rise = thp - tstar(bot - 1, L, par);
          mc = fmin(A / (rise > 1 ? rise : 1.0), mu[bot] - par[P_FLOOR]) * omexp(dt / par[P_TAU_CONV], mix);
          if (par[P_PLUME_ROOT] > 0 && q0 > 0) mc = fmin(mc, mu[bot] * q[bot] / q0);
          /* the mixed boundary layer (bl_q): plumes leave it through the clouds at its top, at most rho a w* with a
           * their share there and w* the free convection's velocity scale (Neggers, Stevens & Neelin 2006) */
          if (blm && par[P_BL_Q] > 0) {
            double l_, f_, d_;
            if (mixed_cloud(L, bot, mu, th[bot], q[bot] + (cloudy ? cl[bot] : 0.0), par, aeps, mix, &l_, &f_, &d_,
                            par[P_CU_CLOUD] > 0)) {
              const double rho = ps / (rd * Tsa), Bk = fmax(0.0, (H + air->vf * cp * Tsa * E) / (rho * cp));
              const double hb = rd * Tsa / grav * log(ps / fmax(ps - grav * mu[bot], 1.0));
              const double ws = cbrt(grav / Tsa * Bk * hb);
              mc = fmin(mc, rho * f_ * ws * dt);
            }
          }
          /* cu_dth: through the cumulus and cloud layers, the air it took in from each (the layer it stops in
           * included) moved by plume_out */
          double cond_cu = 0, det_cu = 0, G_cu = 1;
          const int cu = par[P_CU_DTH] > 0;
          if (mc > 0 && cu)
            plume_out(L, par, mu, th, q, pib, cloudy ? cl : NULL, kd, &mc, th0, q0, thp, qc, Gl, qx, wxo, par[P_L_DET],
                      &cond_cu, &det_cu, &G_cu);
          if (mc > 0) {
            if (!(kd < bot - 1)) G = 1;
            const double me = cu ? mc * (G_cu - 1) : fmin(mc * (G - 1), fmax(0.0, mu[bot - 1] - par[P_FLOOR]));
            const double qe = q[bot - 1], the = th[bot - 1];
            G = cu ? G_cu : 1 + me / mc;
            const double cond_ = cu ? cond_cu : q0 + (G - 1) * qe - G * qc - xs;
            const double cond = cu ? cond_cu : fmax(0.0, cond_);
            const double det = cu ? det_cu : cloudy ? fmin(cond, par[P_L_DET] * G) : 0.0;
            if (flash) flash[i] = storm_flashes(mc / dt, lt, par, air, mix);
            R += mc * (cond - det);
            Rc = mc * (cond - det);
            if (!cu) {
              const double qarr = qc + (cond_ - cond) / G;
              const double Ec = column_energy(L, mu, pib, th, q, par);
              for (int k = 0; k < bot; k++) q[k] = fmax(0.0, q[k] + mc * qx[k] / mu[k]);
              move_mass(L, par, mu, th, q, pib, bot, kd, mc, th0, thp, q0, qarr, cloudy ? cl : NULL);
              if (me > 0) move_mass(L, par, mu, th, q, pib, bot - 1, kd, me, the, thp, qe, qarr, cloudy ? cl : NULL);
              th[kd] += (Ec - column_energy(L, mu, pib, th, q, par)) / (mu[kd] * pib[kd]);
            }
            double Sc = 0;
            if (par[P_ICE_AIR] > 0 && cond > 0 && cond > det) {   /* (its frozen share snows, melting below Tf) */
              Sc = mc * (cond - det) * fmin(1.0, FZ / cond);
              th[kd] += Lf * Sc / (mu[kd] * pib[kd]);
              dgl[kd] += Lf * Sc / dt;
              for (int j = kd + 1; j <= bot && Sc > 0; j++)
                if (base_T(L, j, mu, th, par, mix) > Tf) {
                  th[j] -= Lf * Sc / (mu[j] * pib[j]);
                  dgl[j] -= Lf * Sc / dt;
                  Sc = 0;
                }
            }
            if (!cu) {
              for (int j = kd; j < bot; j++) wxo[j] += mc;
              for (int j = kd; j < bot - 1; j++) wxo[j] += me;
            }
            cin[kd] += mc + me;
            if (cloudy) {
              cl[kd] += mc * det / mu[kd];
              cloud_adjust(L, kd, mu, th, q, cl, pib, dt, par, aeps, mix);
            } else {
              const double rt = mu[kd] * condense_k(L, kd, mu, th, q, pib, par, aeps, mix);
              R += rt;
              Rc += rt;
            }
            if (par[P_DOWNDRAFT] > 0 && bot >= 2 && Rc - Sc > 0) {
              int s_ = bot - 1;
              if (cu && bot >= 3) {
                /* (cu_dth: it sinks from the cumulus layer or the cloud layer above it, whichever has the lower
                 * equivalent potential temperature theta exp(Lv q / (cp T)) at its middle) */
                double pc = 0, te_[2];
                for (int k = 0; k < bot - 2; k++) pc += grav * mu[k];
                for (int m = 0; m < 2; m++) {
                  const int k = bot - 2 + m;
                  const double pm = pc + 0.5 * grav * mu[k], Tk = th[k] * fpow_(pm / p0, kappa, mix);
                  te_[m] = th[k] * exp(Lv * q[k] / (cp * Tk));
                  pc += grav * mu[k];
                }
                if (te_[0] < te_[1]) s_ = bot - 2;
              }
              double pt_ = 0;
              for (int k = 0; k < s_; k++) pt_ += grav * mu[k];
              const double p1 = pt_ + 0.5 * grav * mu[s_];
              const double p2 = s_ == bot - 1 ? pt_ + grav * (mu[s_] + mu[bot]) : pt_ + grav * (mu[s_] + mu[s_ + 1] + mu[bot]);
              double thd_ = th[s_], qd_ = q[s_];
              for (int j = 0; j <= 4; j++) wet_bulb(&thd_, &qd_, p1 + (p2 - p1) * j / 4, par, aeps, mix);
              const double ev = qd_ - q[s_];
              double md = fmin(par[P_DOWNDRAFT] * mc, fmax(0.0, mu[s_] - par[P_FLOOR]));
              if (ev > 0) md = fmin(md, (Rc - Sc) / ev);
              if (md > 0 && thd_ < th[s_]) {
                move_mass(L, par, mu, th, q, pib, s_, bot, md, th[s_], thd_, q[s_], qd_, cloudy ? cl : NULL);
                for (int j = s_; j < bot; j++) wxo[j] -= md;
                const double e_ = md * fmax(0.0, ev);
                R -= e_;
                Rc -= e_;
                dgl[bot] -= Lv * e_ / dt;
              }
            }
            S += Sc;
            Scg = Sc;
          } else {
            mc = 0;
          }
        }
      } else {
      parcel(th0, q0, pi_, par, aeps, mix, &thp, &qc);
      const double d = thp - tstar(bot - 1, L, par);
      if (d > 0) {
        const double dp = 500;
        parcel(th0, q0, pi_ + dp, par, aeps, mix, &thp2, &qc2);
        const double slope = (thp - thp2) / dp;
        const double maxup = mu[bot] - par[P_FLOOR];
        const double want = slope > 0 ? d / slope / grav : maxup;
        mc = fmin(want, maxup) * omexp(dt / par[P_TAU_CONV], mix);
        if (mc > 0) {
          int kd = bot - 1;
          double pj = pi_;
          while (kd > 0) {
            pj -= grav * mu[kd];
            double t2, q2;
            parcel(th0, q0, pj, par, aeps, mix, &t2, &q2);
            if (!(t2 > tstar(kd - 1, L, par))) break;
            thp = t2;
            qc = q2;
            kd--;
          }
          R += mc * (q0 - qc);
          Rc = mc * (q0 - qc);
          move_mass(L, par, mu, th, q, pib, bot, kd, mc, th0, thp, q0, qc, NULL);
          for (int j = kd; j < bot; j++) wxo[j] += mc;
          cin[kd] += mc;
          const double rt = mu[kd] * condense_k(L, kd, mu, th, q, pib, par, aeps, mix);
          R += rt;
          Rc += rt;
        } else {
          mc = 0;
        }
      }
      }
      /* shallow (air.cu:1730-1835, without the cumulus layer): a second plume from the same root (plume_root with
       * root_q), entraining twice the deep one's with ent_ifs, stopping in the cloud layer and leaving its air there;
       * its base mass flux Grant (2001)'s rho a w* (a = shallow), cut to the air and vapor the layers hold. Its
       * condensate leaves as the deep plume's: l_det as cloud, the rest rains, its frozen share as snow */
      if (par[P_SHALLOW] > 0 && par[P_PLUME_DEEP] > 0 && bot >= 2 && ex_r >= 0 && ws_r > 0) {
        double pis = 0;
        for (int k = 0; k < bot; k++) pis += grav * mu[k];
        const double ths0 = th[bot];
        double qs0 = edge_q(L, bot, 0, mu, th, q, 0, par, aeps, mix);
        qs0 += fmin(ex_r, fmax(0.0, qss_r - qs0));
        double As, Gs = 1, FZs = 0, thps, qcs, qxs[OG_LMAX], xss = 0, lts[4], Gls[OG_LMAX];
        int kds;
        plume_ascent(L, mu, th, ths0, qs0, pis, par, air, mix, &As, &kds, &thps, &qcs, q, qxs, lts, &Gs, &FZs, 1, Gls,
                     NULL);
        for (int k = 0; k <= bot; k++) xss += qxs[k];
        if (par[P_CU_DTH] > 0) {
          /* cu_dth: through the cumulus layer and into the cloud layer above it, stopping in either; what it took in
           * from each moved by plume_out; with buoy_v only through the cloud at the mixed layer's top, its share
           * there at most */
          if ((kds == bot - 1 || kds == bot - 2) && As > 0) {
            double ms = par[P_SHALLOW] * rho_r * ws_r * dt, cond, det, Gx;
            if (par[P_BUOY_V] > 0) {
              double l_, f_, d_;
              ms = mixed_cloud(L, bot, mu, th[bot], q[bot] + (cloudy ? cl[bot] : 0.0), par, aeps, mix, &l_, &f_, &d_,
                               par[P_CU_CLOUD] > 0)
                       ? fmin(ms, rho_r * f_ * ws_r * dt)
                       : 0.0;
            }
            if (qs0 > 0) ms = fmin(ms, mu[bot] * q[bot] / qs0);
            plume_out(L, par, mu, th, q, pib, cloudy ? cl : NULL, kds, &ms, ths0, qs0, thps, qcs, Gls, qxs, wxo,
                      par[P_L_DET], &cond, &det, &Gx);
            if (ms > 0) {
              const double rs = ms * (cond - det);
              R += rs;
              Rc += rs;
              if (par[P_ICE_AIR] > 0 && cond > det) {    /* (its frozen share snows, melting below where warm) */
                double Ss = rs * fmin(1.0, FZs / cond);
                th[kds] += Lf * Ss / (mu[kds] * pib[kds]);
                dgl[kds] += Lf * Ss / dt;
                for (int j = kds + 1; j <= bot && Ss > 0; j++)
                  if (base_T(L, j, mu, th, par, mix) > Tf) {
                    th[j] -= Lf * Ss / (mu[j] * pib[j]);
                    dgl[j] -= Lf * Ss / dt;
                    Ss = 0;
                  }
                S += Ss;
                Scg += Ss;
              }
              cin[kds] += ms * Gx;
              if (cloudy) {
                cl[kds] += ms * det / mu[kds];
                cloud_adjust(L, kds, mu, th, q, cl, pib, dt, par, aeps, mix);
              }
              mc += ms;
            }
          }
        } else if (kds == bot - 1 && As > 0) {
          /* its air taken in from the cloud layer (Gs - 1 per kg leaving the mixed layer) leaves it at the layer's
           * vapor and comes back at the plume's: at most the layer's air above its floor, and what it and the mixing
           * take at most half the layer's vapor */
          const int ks = bot - 1;
          const double qe = q[ks];
          double ms = fmin(par[P_SHALLOW] * rho_r * ws_r * dt, mu[bot] - par[P_FLOOR]);
          if (qs0 > 0) ms = fmin(ms, mu[bot] * q[bot] / qs0);
          if (Gs > 1) ms = fmin(ms, (mu[ks] - par[P_FLOOR]) / (Gs - 1));
          for (int k = 0; k < bot; k++) {
            const double take = (k == ks ? (Gs - 1) * qe : 0.0) - qxs[k];
            if (take > 0) ms = fmin(ms, 0.5 * mu[k] * q[k] / take);
          }
          if (ms > 0) {
            const double cond_ = qs0 + (Gs - 1) * qe - Gs * qcs - xss, cond = fmax(0.0, cond_);
            const double det = cloudy ? fmin(cond, par[P_L_DET] * Gs) : 0.0, rs = ms * (cond - det);
            const double qarr = Gs * qcs + (cond_ - cond) - (Gs - 1) * qe;   /* (per kg moved: Gs kg arrive) */
            const double Ec = column_energy(L, mu, pib, th, q, par);
            for (int k = 0; k < bot; k++) q[k] = fmax(0.0, q[k] + ms * qxs[k] / mu[k]);
            move_mass(L, par, mu, th, q, pib, bot, kds, ms, ths0, thps, qs0, qarr, cloudy ? cl : NULL);
            th[kds] += (Ec - column_energy(L, mu, pib, th, q, par)) / (mu[kds] * pib[kds]);
            R += rs;
            Rc += rs;
            if (par[P_ICE_AIR] > 0 && cond > det) {      /* (its frozen share snows, melting in a warm bottom layer) */
              const double Ss = rs * fmin(1.0, FZs / cond);
              th[kds] += Lf * Ss / (mu[kds] * pib[kds]);
              dgl[kds] += Lf * Ss / dt;
              if (base_T(L, bot, mu, th, par, mix) > Tf) {
                th[bot] -= Lf * Ss / (mu[bot] * pib[bot]);
                dgl[bot] -= Lf * Ss / dt;
              } else {
                S += Ss;
                Scg += Ss;
              }
            }
            wxo[kds] += ms;
            cin[kds] += ms;
            if (cloudy) {
              cl[kds] += ms * det / mu[kds];
              cloud_adjust(L, kds, mu, th, q, cl, pib, dt, par, aeps, mix);
            }
            mc += ms;
          }
        }
      }
      /* cu_sub (air.cu:1834-1861): the air the plumes took out of the mixed layer comes back from the cumulus layer,
       * sinking between the clouds, up to the condensation level: the slice from that layer's bottom edge, its theta
       * and vapor as the mixed layer's entrainment's; sub_mix stirs it toward the cumulus layer's mean */
      if (par[P_CU_SUB] > 0 && par[P_BL_LCL] > 0 && blm && bot >= 2) {
        const int ka = bot - 1;
        const double m = fmin(fmin(mu_pre - mu[bot], lcl_mass(L, mu, th, q, par, aeps, mix) - mu[bot]),
                              mu[ka] - par[P_FLOOR]);
        if (m > 0) {
          const double e = tstar(ka, L, par), fm = fmin(1.0, m / mu[ka]);
          const int spread = th[ka] < tstar(ka - 1, L, par) && th[ka] > e;
          double thc = spread ? e + fm * (th[ka] - e) : th[ka];
          double qx = par[P_EDGE_VAP] > 0 ? edge_q(L, ka, 0, mu, th, q, m, par, aeps, mix) : q[ka];
          if (par[P_SUB_MIX] > 0) {
            thc += par[P_SUB_MIX] * (th[ka] - thc);
            qx += par[P_SUB_MIX] * (q[ka] - qx);
          }
          move_mass(L, par, mu, th, q, pib, ka, bot, m, thc, thc, qx, qx, cloudy ? cl : NULL);
          wxo[ka] -= m;
        }
      }
    }
    /* the plumes' rain evaporates as it falls below their base through the bottom layer's clear air (revap: on the
     * convection's area revap_area; else on conv_revap of the cell; with downdrafts they evaporate it instead) */
    if (cloudy && (par[P_REVAP] > 0 ? Rc - Scg > 0 : par[P_CONV_REVAP] > 0 && !(par[P_DOWNDRAFT] > 0) && Rc > 0) && L > 1) {
      double qs, dqs;
      layer_cap(L, bot, mu, th[bot], par, aeps, mix, &qs, &dqs);
      const double rc_ = par[P_REVAP] > 0 ? (isl ? par[P_REVAP_RHL] : par[P_REVAP_RHS]) : 1.0;
      if (rc_ * qs > q[bot]) {
        double ev;
        if (par[P_REVAP] > 0) {
          const double Fc = (Rc - Scg) / dt;
          double pt_ = 0;
          for (int k = 0; k < bot; k++) pt_ += grav * mu[k];
          const double pm = pt_ + 0.5 * grav * mu[bot], psf = pt_ + grav * mu[bot];
          ev = par[P_REVAP_AREA] * 5.44e-4 * (rc_ * qs - q[bot]) *
               pow(sqrt(pm / psf) / 5.09e-3 * Fc / par[P_REVAP_AREA], 0.577) * mu[bot];
          ev = fmin(ev, fmin(Fc, (rc_ * qs - q[bot]) * mu[bot] / dt / (1 + Lv / pib[bot] * dqs)));
        } else {
          const double Fc = Rc / dt;
          ev = par[P_K_REVAP] * (1 - q[bot] / qs) * sqrt(par[P_CONV_REVAP] * Fc) * mu[bot];
          ev = fmin(ev, fmin(Fc, (qs - q[bot]) * mu[bot] / dt / (1 + Lv / pib[bot] * dqs)));
        }
        q[bot] += ev * dt / mu[bot];
        th[bot] -= Lv * ev * dt / (mu[bot] * pib[bot]);
        dgl[bot] -= Lv * ev;
        R -= ev * dt;
        Rc -= ev * dt;
      }
    }
    if (conv) conv[i] = mc;
    for (int j = 0; j + 1 < L; j++) wx[(size_t)j * N + i] = wxo[j];
    if (xfer)
      for (int g = 0; g < OG_GASN; g++)
        for (int k = 0; k < L; k++) xfer[((size_t)g * L + k) * N + i] = xfl[g * OG_LMAX + k];
    xf_col = NULL;
    if (dgo) {                                       /* rows: Q (L), latent heating (L), gross up and down (2 (L - 1)),
                                                      * convective mass (L) */
      int r = 0;
      for (int k = 0; k < L; k++) dgo[(size_t)(r++) * N + i] = Q[k];
      for (int k = 0; k < L; k++) dgo[(size_t)(r++) * N + i] = dgl[k];
      for (int j = 0; j + 1 < L; j++) {
        dgo[(size_t)(r++) * N + i] = gross[2 * j];
        dgo[(size_t)(r++) * N + i] = gross[2 * j + 1];
      }
      for (int k = 0; k < L; k++) dgo[(size_t)(r++) * N + i] = cin[k];
    }
    rain[i] = R / dt;
    if (!isl && S > 0) {                             /* snow into the sea melts with the slab's heat */
      Ts[i] -= Lf * S / Cw;
      if (qocean) qocean[i] -= Lf * S / dt;
    }
    if (isl) {
      /* the ice's melt runs into the ponds, from the snow first. On ice the skin decides: below freezing rain and
       * snow freeze as they land, a pond as fast as the surface can lose the heat (to the cold ice, the cold air, the
       * sky; and into snow without firn at 10 mm/h), the freezing heat into the air touching the ground; at freezing
       * rain runs on and the snow melts from that air. Bare ground: snow lies on frozen ground, its freezing (the
       * rain's share) warming it; else rain runs to the ponds. A pond on frozen ground freezes up to its heat room;
       * on thawed ground it soaks in at k_inf toward the bucket's capacity, and a full bucket springs. */
      pw += ice0 - m_ice;
      double sn = snow[i] - (ice0 - m_ice);
      if (skin) {
        if (tsk < Tf) {
          const double room = (par[P_K_ICE] / dice * (Tf - tsk) + hs * fmax(0.0, Tf - Tsa) +
                               fmax(0.0, eps_s * sigma * Tf * Tf * Tf * Tf - ldown - swS)) * dt / Lf +
                              (sn > 0 && !fon ? 10.0 / 3600 * dt : 0.0);
          const double mp = fmin(pw > 0 ? pw : 0.0, room), m = R + mp;
          m_ice += m;
          sn = (sn > 0 ? sn : 0.0) + R;
          pw -= mp;
          for (int k = 0; k <= bot; k++) th[k] += Lf * (m - S) * (wbl[k] / mb) / (mu[k] * pib[k]);
        } else {
          pw += R;
          for (int k = 0; k <= bot; k++) th[k] -= Lf * S * (wbl[k] / mb) / (mu[k] * pib[k]);
        }
        if (fon && pw > 0 && Tfn < Tf) {             /* melt and rain soak into the cold firn and refreeze */
          const double mr = fmin(pw, Cf * (Tf - Tfn) / Lf);
          pw -= mr;
          m_ice += mr;
          sn = (sn > 0 ? sn : 0.0) + mr;
          Tfn += Lf * mr / Cf;
        }
      } else if (Ts[i] < Tf) {
        m_ice += R;
        sn = (sn > 0 ? sn : 0.0) + R;
        Ts[i] += Lf * (R - S) / Cw;
      } else {
        pw += R;
        Ts[i] -= Lf * S / Cw;
      }
      if (!skin && Ts[i] < Tf && pw > 0) {
        const double room = (Tf - Ts[i]) * Cw / Lf, m = pw < room ? pw : room;
        pw -= m;
        m_ice += m;
        Ts[i] += Lf * m / Cw;
      }
      ice[i] = m_ice;
      snow[i] = fmax(0.0, fmin(sn, m_ice));
      if (!skin && Ts[i] >= Tf) {
        const double room = par[P_K_INF] * dt * (1 - wet / soilcap[i]);
        const double m = pw < room ? pw : room;
        if (m > 0) {
          pw -= m;
          wet += m;
        }
      }
      if (wet > soilcap[i]) {
        pw += wet - soilcap[i];
        wet = soilcap[i];
      }
      soil[i] = wet;
      pond[i] = pw;
    }
    if (rainp) {
      rainp[i] = Rk[0] / dt;
      rainp[(size_t)N + i] = Rk[1] / dt;
      rainp[(size_t)2 * N + i] = Rc / dt;
      rainp[(size_t)3 * N + i] = S / dt;
    }
    /* a layer emptied to its floor keeps a theta within its edges: its heat beyond them goes to the nearest layer
     * with mass (below first), the column's energy unchanged */
    if (blm)
      for (int k = 1; k < bot; k++) {
        if (mu[k] > 2 * par[P_FLOOR]) continue;
        const double lo = tstar(k, L, par), hi = tstar(k - 1, L, par), tn = fmin(fmax(th[k], lo), hi);
        if (tn == th[k]) continue;
        int n = -1;
        for (int j = k + 1; j < L && n < 0; j++)
          if (mu[j] > 2 * par[P_FLOOR]) n = j;
        for (int j = k - 1; j >= 0 && n < 0; j--)
          if (mu[j] > 2 * par[P_FLOOR]) n = j;
        if (n < 0) continue;
        th[n] += mu[k] * pib[k] * (th[k] - tn) / (mu[n] * pib[n]);
        th[k] = tn;
      }
    const double E1 = column_energy(L, mu, pib, th, q, par) + Cw * Ts[i] - Lf * m_ice + Cf * (Tfn - Tf);
    if (Cf > 0) firn[i] = Tfn;
    for (int k = 0; k < L; k++) {
      theta[(size_t)k * N + i] = th[k];
      if (moist) qv[(size_t)k * N + i] = q[k];
      if (cloudy) cloudw[(size_t)k * N + i] = cl[k];
    }
    stored[i] = E1 - E0;
#undef SOILF
  }
}
