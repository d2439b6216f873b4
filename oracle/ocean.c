/* The ocean's column processes on the CPU -- TEST INFRASTRUCTURE ONLY (the oracle; nothing in the product links it).
 *
 * A scalar restatement of the reference's ocean mixing kernels (src/insilica/kernels/ocean.cu) in their operation
 * order, one column at a time, for the tests' parity checks of the HIP library and for the CPU baseline:
 *   og_ocn_overturn  -- ocn_overturn / adjust (ocean.cu:55-72, 231-237): convective adjustment of T and S;
 *   og_ocn_column    -- ocn_column (ocean.cu:117-213): Kraus-Turner entrainment or shoaling of the mixed layer, the
 *                       base's shear mixing (ri_b), upwelling to h_ent, reach through emptied layers, the interior
 *                       interfaces' deep-water formation (dwf) and diapycnal mixing (kappa_d), detrainment into the
 *                       class of its own density (det_class), into wx (L - 1, N);
 *   og_ocn_xch_salt  -- ocn_xch_salt (ocean.cu:352-381): salt moved with wx (upwind; seq: passing order);
 *   og_ocn_pen       -- ocn_pen (ocean.cu:387-407): the sunlight passing the mixed layer's base heats the layers below;
 *                       the mixing's flux sums;
 *   og_ocn_fresh     -- ocn_fresh (ocean.cu:245-261): rain, evaporation and sea ice freshen or salt the mixed layer;
 *   og_ocn_diffuse   -- ocn_diffuse (ocean.cu:268-292): mixing of T and S along each layer.
 * Layout: layer-major (L, N) arrays, as the reference's.
 *
 * Pinned in tests/test_oracle_ocean.py by the budgets these processes keep (heat, salt and volume per column and in
 * total; salt against the fresh water), closed forms (the mixed layer's Kraus-Turner entrainment rate P / (h db) dt,
 * the Monin-Obukhov depth where the energy balances, the shear depth sqrt(ri_b du^2 h / db), Paulson & Simpson's
 * light), and the adjustment's stable end state found by brute force.
 */
#include <math.h>
#include <stddef.h>

#include "oracle.h"

#define OLMAX 8

typedef struct {
  double g, a, a2, beta;
} og_eos;

/* buoyancy b = g (alpha th + alpha2 th^2 / 2 - beta S), th = T - 273.15 (ocean.cu:47-50) */
static double buoy(double T, double S, const og_eos* q) {
  const double th = T - 273.15;
  return q->g * (q->a * th + 0.5 * q->a2 * th * th - q->beta * S);
}
static double expand(double T, const og_eos* q) { return q->a + q->a2 * (T - 273.15); }

/* adjust (ocean.cu:55-72): blocks of mixed layers from the top down; a block denser than the one below merges with
 * it (heat and salt mixed by thickness), until the column is stable */
static void adjust(int N, int L, int i, const double* h, double* T, double* S, const og_eos* q) {
  double hb[OLMAX], Tb[OLMAX], Sb[OLMAX];
  int first[OLMAX], nb = 0;
  for (int k = 0; k < L; k++) {
    hb[nb] = h[(size_t)k * N + i];
    Tb[nb] = T[(size_t)k * N + i];
    Sb[nb] = S[(size_t)k * N + i];
    first[nb] = k;
    nb++;
    while (nb > 1 && buoy(Tb[nb - 2], Sb[nb - 2], q) < buoy(Tb[nb - 1], Sb[nb - 1], q)) {
      const double hs = hb[nb - 2] + hb[nb - 1];
      Tb[nb - 2] = (hb[nb - 2] * Tb[nb - 2] + hb[nb - 1] * Tb[nb - 1]) / hs;
      Sb[nb - 2] = (hb[nb - 2] * Sb[nb - 2] + hb[nb - 1] * Sb[nb - 1]) / hs;
      hb[nb - 2] = hs;
      nb--;
    }
  }
  for (int j = 0; j < nb; j++) {
    const int k1 = j + 1 < nb ? first[j + 1] : L;
    for (int k = first[j]; k < k1; k++) {
      T[(size_t)k * N + i] = Tb[j];
      S[(size_t)k * N + i] = Sb[j];
    }
  }
}

/* Paulson & Simpson's light passing depth z and its integral from the surface (ocean.cu:75-79) */
static double sun_f(double z, double R, double z1, double z2) { return R * exp(-z / z1) + (1 - R) * exp(-z / z2); }
static double sun_int(double z, double R, double z1, double z2) {
  return -(R * z1 * expm1(-z / z1) + (1 - R) * z2 * expm1(-z / z2));
}
/* the buoyancy flux a mixed layer of depth h works against, sunlight by where it is absorbed (ocean.cu:84-86) */
static double kt_buoy(double h, double Bns, double Bsw, double R, double z1, double z2) {
  return Bns + Bsw * (1 + sun_f(h, R, z1, z2) - 2 * sun_int(h, R, z1, z2) / h);
}
/* the turbulent energy at the base: 2 m u*^3 less the work against heating or plus n of the convection's (:90-92) */
static double kt_energy(double h, double us3, double B, double m_kt, double n_kt) {
  return 2 * m_kt * us3 - 0.5 * h * ((1 + n_kt) * B + (1 - n_kt) * fabs(B));
}

void og_ocn_overturn(int N, int L, const double* h, double* T, double* S, double grav, double alpha, double alpha2,
                     double beta, const double* land) {
  const og_eos q = {grav, alpha, alpha2, beta};
#pragma omp parallel for schedule(static)
  for (int i = 0; i < N; i++) {
    if (land && land[i] > 0.5) continue;
    adjust(N, L, i, h, T, S, &q);
  }
}

/* ocn_column (ocean.cu:117-213), one column */
static void column(int N, int L, int i, const double* h, double* T, double* S, const double* taum, const double* qns,
                   const double* qfw, const double* qsun, const og_eos* q, double S_ref, double alpha_ref, double rho_w,
                   double cw, double m_kt, double n_kt, double hmin, double hmax, double dTmin, double R, double z1,
                   double z2, double dt, double* wx, const double* land, double kappa_d, double dwf,
                   const double* shear, double ri_b, double reach, double h_ent, const double* Tc, double det_class,
                   const double* Sc) {
  if (land && land[i] > 0.5) {
    for (int k = 0; k < L - 1; k++) wx[(size_t)k * N + i] = 0;
    return;
  }
  adjust(N, L, i, h, T, S, q);
  const double h0 = h[i], h1 = h[(size_t)N + i];
  int kb = 1;
  if (reach > 0)
    while (kb < L - 1 && h[(size_t)kb * N + i] - hmin < 1) kb++;
  const double T0 = T[i], T1 = T[(size_t)kb * N + i];
  const double us = sqrt(taum[i] / rho_w);
  const double us3 = us * us * us;
  const double ga = q->g * expand(T0, q);
  const double Bns = ga * qns[i] / (rho_w * cw) + q->g * q->beta * S_ref * qfw[i] / rho_w;
  const double Bsw = ga * qsun[i] / (rho_w * cw);
  const double P = kt_energy(h0, us3, kt_buoy(h0, Bns, Bsw, R, z1, z2), m_kt, n_kt);
  const double db = buoy(T0, S[i], q) - buoy(T1, S[(size_t)kb * N + i], q);
  double dm = 0;
  if (P > 0) {
    if (db > q->g * alpha_ref * dTmin) dm = P / (h0 * db) * dt;
  } else {
    double lo = hmin, hi = h0;
    if (kt_energy(lo, us3, kt_buoy(lo, Bns, Bsw, R, z1, z2), m_kt, n_kt) > 0)
      for (int it = 0; it < 50; it++) {
        const double mid = 0.5 * (lo + hi);
        if (kt_energy(mid, us3, kt_buoy(mid, Bns, Bsw, R, z1, z2), m_kt, n_kt) > 0)
          lo = mid;
        else
          hi = mid;
      }
    dm = lo - h0;
  }
  if (shear && ri_b > 0 && db > 0) {
    const double du = shear[i];
    const double hn = sqrt(ri_b * du * du * h0 / db);
    if (hn - h0 > dm) dm = hn - h0;
  }
  if (h_ent > 0 && h0 + dm < h_ent) dm = h_ent - h0;
  double up_max = h1 - hmin;
  if (reach > 0) {
    up_max = 0;
    for (int k = 1; k <= kb; k++) up_max += fmax(h[(size_t)k * N + i] - hmin, 0.0);
  }
  if (hmax - h0 < up_max) up_max = hmax - h0;
  const double dn_max = h0 - hmin;
  if (dm > up_max) dm = up_max > 0 ? up_max : 0.0;
  if (-dm > dn_max) dm = dn_max > 0 ? -dn_max : 0.0;
  wx[i] = dm;
  double add[OLMAX + 1];
  for (int k = 0; k <= OLMAX; k++) add[k] = 0;
  if (reach > 0 && dm > 0) {
    double rem = dm - fmax(h1 - hmin, 0.0);
    for (int k = 1; k < L - 1 && rem > 0; k++) {
      add[k] = rem;
      rem -= fmax(h[(size_t)(k + 1) * N + i] - hmin, 0.0);
    }
  }
  double gave = dm > 0 ? dm - add[1] : 0.0;
  for (int k = 1; k < L - 1; k++) {
    const double hk = h[(size_t)k * N + i], hb = h[(size_t)(k + 1) * N + i];
    double v = 0;
    if (dwf > 0 && Bns < 0 && add[k] <= 0 &&
        buoy(T[(size_t)k * N + i], S[(size_t)k * N + i], q) <=
            buoy(T[(size_t)(k + 1) * N + i], S[(size_t)(k + 1) * N + i], q))
      v = -fmax(hk - hmin - gave, 0.0);
    else if (kappa_d > 0)
      v = fmin(2 * kappa_d / (hk + hb) * dt, fmax(hb - hmin - (add[k] - add[k + 1]), 0.0));
    wx[(size_t)k * N + i] = v + add[k];
    gave = (v > 0 ? v : 0.0) + (add[k] - add[k + 1]);
  }
  if (Tc && det_class > 0 && dm < 0) {
    int ks = 1;
    const double b0 = buoy(T0, S[i], q);
    for (int k = 2; k < L; k++)
      if (b0 <= 0.5 * (buoy(Tc[k - 1], Sc ? Sc[k - 1] : S_ref, q) + buoy(Tc[k], Sc ? Sc[k] : S_ref, q))) ks = k;
    for (int k = 1; k < ks; k++) wx[(size_t)k * N + i] += dm;
  }
}

void og_ocn_column(int N, int L, const double* h, double* T, double* S, const double* taum, const double* qns,
                   const double* qfw, const double* qsun, double grav, double alpha, double alpha2, double beta,
                   double S_ref, double alpha_ref, double rho_w, double cw, double m_kt, double n_kt, double hmin,
                   double hmax, double dTmin, double R, double z1, double z2, double dt, double* wx,
                   const double* land, double kappa_d, double dwf, const double* shear, double ri_b, double reach,
                   double h_ent, const double* Tc, double det_class, const double* Sc) {
  const og_eos q = {grav, alpha, alpha2, beta};
#pragma omp parallel for schedule(static)
  for (int i = 0; i < N; i++)
    column(N, L, i, h, T, S, taum, qns, qfw, qsun, &q, S_ref, alpha_ref, rho_w, cw, m_kt, n_kt, hmin, hmax, dTmin, R,
           z1, z2, dt, wx, land, kappa_d, dwf, shear, ri_b, reach, h_ent, Tc, det_class, Sc);
}

void og_ocn_xch_salt(int N, int L, const double* wx, const double* h, double* S, int seq) {
#pragma omp parallel for schedule(static)
  for (int i = 0; i < N; i++) {
    double hs[OLMAX], tr[OLMAX];
    for (int k = 0; k < L; k++) {
      hs[k] = h[(size_t)k * N + i];
      tr[k] = hs[k] * S[(size_t)k * N + i];
    }
    if (seq) {
      for (int pass = 0; pass < 2; pass++)
        for (int n = 0; n + 1 < L; n++) {
          const int k = pass == 0 ? L - 2 - n : n;
          const double v = wx[(size_t)k * N + i];
          if (pass == 0 ? !(v > 0) : !(v < 0)) continue;
          const int d = v > 0 ? k + 1 : k;
          const double s = tr[d] / hs[d];
          tr[k] += v * s;
          tr[k + 1] -= v * s;
          hs[k] += v;
          hs[k + 1] -= v;
        }
    } else {
      for (int k = 0; k + 1 < L; k++) {
        const double v = wx[(size_t)k * N + i];
        const double s = v > 0 ? S[(size_t)(k + 1) * N + i] : S[(size_t)k * N + i];
        tr[k] += v * s;
        tr[k + 1] -= v * s;
        hs[k] += v;
        hs[k + 1] -= v;
      }
    }
    for (int k = 0; k < L; k++) S[(size_t)k * N + i] = tr[k] / hs[k];
  }
}

void og_ocn_pen(int N, int L, const double* h, double* T, const double* qocean, const double* qsw, double rho_cw,
                double R, double z1, double z2, double dt, double* qns_acc, double* sun_acc, const double* land,
                const double* up) {
#pragma omp parallel for schedule(static)
  for (int i = 0; i < N; i++) {
    if (land && land[i] > 0.5) continue;
    qns_acc[i] += (qocean[i] - qsw[i]) * dt;
    sun_acc[i] += qsw[i] * dt;
    if (up && up[i] > 0) continue;
    double z = h[i];
    double f = sun_f(z, R, z1, z2);
    const double pen = qsw[i] * f;
    T[i] -= pen * dt / (rho_cw * h[i]);
    for (int k = 1; k < L; k++) {
      const double hk = h[(size_t)k * N + i];
      z += hk;
      const double fb = k < L - 1 ? sun_f(z, R, z1, z2) : 0.0;
      T[(size_t)k * N + i] += qsw[i] * (f - fb) * dt / (rho_cw * hk);
      f = fb;
    }
  }
}

void og_ocn_fresh(int N, const double* h, double* S, const double* evap, const double* rain, const double* ice,
                  double* ice_prev, double S_ref, double rho_w, double dt, double* fw_acc, const double* land,
                  double* mouth, double fw_local, double* hw, double* M, const double* area) {
#pragma omp parallel for schedule(static)
  for (int i = 0; i < N; i++) {
    if (land && land[i] > 0.5) continue;
    double fw = (rain[i] - evap[i]) * dt - (ice[i] - ice_prev[i]);
    if (mouth) {
      fw += mouth[i];
      mouth[i] = 0;
    }
    if (fw_local > 0) {
      const double h1 = fmax(h[i] + fw / rho_w, fmin(h[i], 10.0));
      S[i] *= h[i] / h1;
      hw[i] = h1;
      M[i] = h1 * area[i];
    } else
      S[i] -= S_ref * fw / (rho_w * h[i]);
    ice_prev[i] = ice[i];
    fw_acc[i] += fw;
  }
}

void og_ocn_diffuse(const og_grid* g, int L, const double* h, const double* T, const double* S, double* Tn,
                    double* Sn, const double* wall, double kappa_h, double dt) {
  const int N = g->N;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < N; i++) {
    for (int k = 0; k < L; k++) {
      const size_t a = (size_t)k * N;
      const double hi = h[a + i], Ti = T[a + i], Si = S[a + i];
      double fT = 0, fS = 0;
      for (int m = 0; m < 6; m++) {
        const int e = g->cellEdge[i * 6 + m];
        if (e < 0) break;
        if (wall && wall[e] == 0) continue;
        const int j = g->eI[e] == i ? g->eJ[e] : g->eI[e];
        const double he = fmin(hi, h[a + j]), c = kappa_h * he * g->le[e] / g->de[e] * dt;
        fT += c * (T[a + j] - Ti);
        fS += c * (S[a + j] - Si);
      }
      Tn[a + i] = Ti + fT / (hi * g->areaCell[i]);
      Sn[a + i] = Si + fS / (hi * g->areaCell[i]);
    }
  }
}
