/* The CPU oracle of the Moon's and the Sun's tides -- TEST INFRASTRUCTURE ONLY.
 *
 * A scalar C restatement of the reference's kernels/tides.cu (kepler :26-31, bodies :35-56, tide_pot :58-71,
 * tide_mix :85-122, tide_itd :128-144) as Tides.force and Tides.mix sequence them (tides.py:70-87),
 * FP64 in the reference's operation order. The tests use it as the checker of the HIP library's isw_tide_force and
 * isw_tide_mix.
 *
 * tp (19): tilt, year_s, day_s, ecc, lam_peri, sun_gm, sun_a, moon_gm, moon_a, moon_e, moon_i, moon_T, node_T,
 * aps_T, moon_L0, node0, aps0, love, radius (the reference's sky_par, tides.py:42-51).
 */
#include <math.h>
#include <stddef.h>

#define TPI 3.141592653589793

/* Kepler's equation E - e sin E = M by six Newton steps from E = M; returns the true anomaly, *r = 1 - e cos E. */
static double kepler(double M, double e, double* r) {
  double E = M;
  for (int it = 0; it < 6; it++) E -= (E - e * sin(E) - M) / (1 - e * cos(E));
  *r = 1 - e * cos(E);
  return 2 * atan2(sqrt(1 + e) * sin(E / 2), sqrt(1 - e) * cos(E / 2));
}

/* The Sun's and the Moon's declinations and right ascensions at time t, the day's turn, and their distances in units
 * of their semi-major axes. */
static void bodies(const double* tp, double t, double* decs, double* ras, double* rot, double* rs, double* decm,
                   double* ram, double* rm) {
  const double tilt = tp[0], year_s = tp[1], day_s = tp[2], ecc = tp[3], lam_peri = tp[4];
  double ls;
  *rs = 1;
  if (!(ecc > 0)) {
    ls = 2 * TPI * t / year_s;
  } else {
    const double nu0 = -lam_peri, E0 = 2 * atan(sqrt((1 - ecc) / (1 + ecc)) * tan(nu0 / 2));
    ls = kepler(E0 - ecc * sin(E0) + 2 * TPI * t / year_s, ecc, rs) + lam_peri;
  }
  *ras = atan2(cos(tilt) * sin(ls), cos(ls));
  *decs = asin(sin(tilt) * sin(ls));
  *rot = 2 * TPI * fmod(t, day_s) / day_s;
  const double e = tp[9], inc = tp[10];
  const double peri = tp[16] + 2 * TPI * t / tp[13], node = tp[15] + 2 * TPI * t / tp[12];
  const double lo = kepler(tp[14] - peri + 2 * TPI * t / tp[11], e, rm) + peri;
  const double u = lo - node;
  const double bet = asin(sin(inc) * sin(u)), lam = node + atan2(cos(inc) * sin(u), cos(u));
  *decm = asin(sin(bet) * cos(tilt) + cos(bet) * sin(tilt) * sin(lam));
  *ram = atan2(sin(lam) * cos(tilt) - tan(bet) * sin(tilt), cos(lam));
}

/* The tide-generating potential V (N) at time t - back (tide_pot). */
void og_tide_pot(int N, const double* tp, double t, double back, const double* pos, double* V) {
  const double tt = t - back;
  double decs, ras, rot, rs, decm, ram, rm;
  bodies(tp, tt, &decs, &ras, &rot, &rs, &decm, &ram, &rm);
  const double ls_e = -rot, lm_e = -rot + (ram - ras);
  const double R2 = tp[18] * tp[18];
  const double as = tp[6] * rs, am = tp[8] * rm;
  for (int i = 0; i < N; i++) {
    const double x = pos[3 * i], y = pos[3 * i + 1], z = pos[3 * i + 2];
    const double cs = x * cos(decs) * cos(ls_e) + y * cos(decs) * sin(ls_e) + z * sin(decs);
    const double cm = x * cos(decm) * cos(lm_e) + y * cos(decm) * sin(lm_e) + z * sin(decm);
    V[i] = tp[17] * (tp[5] * R2 / (as * as * as) * (3 * cs * cs - 1) / 2 + tp[7] * R2 / (am * am * am) * (3 * cm * cm - 1) / 2);
  }
}

/* Tidal mixing after an ocean step (tide_mix): the bottom stress's dissipation and the mixing it drives. */
void og_tide_mix(int N, int E, int L, const double* h, const double* u, double* T, double* S, const int* cellEdge,
                 const double* le, const double* de, const double* area, const double* land, double c_bot, double rho,
                 double gamma, double dt, double grav, double alpha, double alpha2, double beta, double* dsum,
                 double cd, const double* cbot, double zeta) {
#pragma omp parallel for schedule(static)
  for (int i = 0; i < N; i++) {
    if (land && land[i] > 0.5) continue;
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
      const double db = grav * (alpha * (t1 - t2) + 0.5 * alpha2 * (t1 * t1 - t2 * t2) - beta * (S[a] - S[b]));
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
}

/* The internal tide's drag coefficient per edge (tide_itd): cbot (E, m/s). nbot (N) or NULL. */
void og_tide_itd(int N, int E, int L, const double* h, const double* T, const double* S, const double* h2,
                 const int* eI, const int* eJ, double kappa, double grav, double alpha, double alpha2, double beta,
                 double* cbot, const double* nbot) {
#pragma omp parallel for schedule(static)
  for (int e = 0; e < E; e++) {
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
      const double db = grav * (alpha * (t1 - t2) + 0.5 * alpha2 * (t1 * t1 - t2 * t2) - beta * (S[a] - S[b]));
      const double dz = 0.5 * (h[a] + h[b]);
      if (db > 0 && dz > 0) c += 0.5 * kappa * h2[i] * sqrt(db / dz);
    }
    cbot[e] = 0.5 * c;
  }
}
