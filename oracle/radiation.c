/* CPU oracle of the k-table radiation -- TEST INFRASTRUCTURE ONLY (parity checker and CPU baseline; never linked into
 * the product).
 *
 * A scalar restatement of the reference's double-precision longwave and shortwave (src/insilica/kernels/air.cu:
 * lw_setup 2145-2229, lw_column 2231-2348, sw_layer 2358-2382, sw_column 2384-2511, with the helpers they call:
 * layer_exner 129-140, qsat/qsat_a 160-171, sun_lon 174-185), in the reference's operation order. The reference's
 * float32 points are kept where its double build has them: the tables, the sub-levels (sub, float32), the table
 * positions clamped through fminf/fmaxf (which round a double argument to float), the differences of the float ozone
 * profile, and the per-g-point results (red[][], float) that the FP64 sum over g-points adds up.
 * Built with gcc -O2 -ffp-contract=off (no fused multiply-adds, as the reference's --fmad=false build).
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define PI_ 3.14159265358979323846
#define NV 15

/* The reference's PAR_ORDER positions of the parameters these kernels read (atmosphere.py:36-43, air.cu:118-127). */
enum { P_EPS_S = 4, P_GRAV = 13, P_CP = 14, P_KAPPA = 15, P_P0 = 16, P_TILT = 17, P_YEAR = 18, P_DAY = 19,
       P_CLOUD_G = 59, P_BL_Q = 89, P_EDGE_K = 101, P_ICE_AIR = 102, P_LW_REFL = 124, P_ECC = 126, P_LAM_PERI = 127 };

/* The layers' pressures at their lower edges and mass-mean Exner functions times cp (air.cu:129-140). */
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

/* Tetens over water (air.cu:160-162) and the ice blend below freezing (air.cu:165-171). */
static double qsat_w(double T, double p, double eps) { return eps * 611.0 * exp(17.27 * (T - 273.15) / (T - 35.85)) / p; }

static double qsat_a(double T, double p, double ice, double eps) {
  const double w = qsat_w(T, p, eps);
  if (!(ice > 0) || T >= 273.15) return w;
  const double l = fmin(1.0, fmax(0.0, (T - 253.15) / 20));
  const double si = eps * 611.0 * exp(21.875 * (T - 273.16) / (T - 7.66)) / p;
  return l * w + (1 - l) * si;
}

/* The sun's longitude at time t and its distance factor (air.cu:174-185): Kepler's equation by five Newton steps. */
double og_sun_lon(double t, const double* par, double* fd) {
  const double ecc = par[P_ECC], year = par[P_YEAR];
  if (!(ecc > 0)) {
    *fd = 1;
    return 2 * PI_ * t / year;
  }
  const double e = ecc, nu0 = -par[P_LAM_PERI];
  const double E0 = 2 * atan(sqrt((1 - e) / (1 + e)) * tan(nu0 / 2));
  const double M = E0 - e * sin(E0) + 2 * PI_ * t / year;
  double E = M;
  for (int it = 0; it < 5; it++) E -= (E - e * sin(E) - M) / (1 - e * cos(E));
  const double nu = 2 * atan2(sqrt(1 + e) * sin(E / 2), sqrt(1 - e) * cos(E / 2)), r = 1 - e * cos(E);
  *fd = 1 / (r * r);
  return nu + par[P_LAM_PERI];
}

/* A table position clamped to [0, n - 1 - 1e-4] as the reference writes it, fminf(fmaxf(x, 0.f), (n - 1) - 1e-4f):
 * both calls take float arguments, so the double position is rounded to float. */
static float clampf_pos(double x, int n) {
  return fminf(fmaxf((float)x, 0.f), (float)((double)(n - 1) - (double)1e-4f));
}

/* ---- lw_setup (air.cu:2145-2229): one (column, layer) at a time ---- */
void og_lw_setup(int N, int L, const double* par, const og_lwt* t, const og_air* air, double co2, double n2o,
                 double ch4, const double* gx, const double* mu_in, const double* theta, const double* qv,
                 const double* cloudo, double hs, int tmode, float* sub) {
  const double cp = par[P_CP], kappa = par[P_KAPPA], p0 = par[P_P0], grav = par[P_GRAV];
  const double bl_q = par[P_BL_Q], ice_air = par[P_ICE_AIR];
  const int edge_k = (int)par[P_EDGE_K];
  static const double FT[5] = {0, 0.05, 0.2, 0.5, 1}, FE[5] = {0, 0.25, 0.5, 0.75, 1};
#pragma omp parallel for schedule(static)
  for (int i = 0; i < N; i++) {
    double mu[OG_LMAX], th[OG_LMAX], q[OG_LMAX], pb[OG_LMAX], pib[OG_LMAX];
    for (int j = 0; j < L; j++) {
      mu[j] = mu_in[(size_t)j * N + i];
      th[j] = theta[(size_t)j * N + i];
      q[j] = qv ? fmax(qv[(size_t)j * N + i], 0.0) : 0.0;
    }
    layer_exner(L, mu, par, pb, pib);
#define W(s, v) sub[((size_t)(s) * NV + (v)) * N + i]
    for (int k = 0; k < L; k++) {
      const double pt = k ? pb[k - 1] : 0.0, pbk = pb[k];
      const int strat = k < edge_k && L >= 3;          /* above the 380 K edge: isothermal, vapor well mixed */
      const double Tiso = pib[k] * th[k] / cp;
      const double nT = tmode == 1 ? kappa : 0.190;
      const double Cn = Tiso * (nT + 1) * (pbk - pt) / (pow(pbk, nT + 1) - pow(pt, nT + 1));
#define TOF(p) ((tmode == 0 || strat) ? Tiso : Cn * pow((p), nT))
      const double fcl = cloudo ? cloudo[(size_t)k * N + i] : 0.0;
      const double tcl = cloudo ? cloudo[(size_t)(L + 3 + k) * N + i] : 0.0;
      double qs[4], qm = 0, uas[4], pe[5];
      for (int j = 0; j <= 4; j++) pe[j] = pt + (k == 0 ? FT[j] : FE[j]) * (pbk - pt);
      for (int j = 0; j < 4; j++) {
        const int s = k * 4 + j;
        const double a = pe[j], b = pe[j + 1], pm = a > 0 ? (b - a) / log(b / a) : 0.5 * b;
        const double Tm = TOF(0.5 * (a + b));
        W(s, 0) = (float)TOF(a);
        W(s, 1) = (float)TOF(b);
        const double Tq = th[k] * pow(0.5 * (a + b) / p0, kappa);
        qs[j] = strat || (bl_q > 0 && k == L - 1 && k > 0) ? 1.0 : (Tq < 160 ? 0.0 : qsat_a(Tq, pm, ice_air, air->eps));
        qm += qs[j] * ((b - a) / (pbk - pt));
        const double x = (t->lp0 - log(pm * air->broad)) / t->dlp, y = (Tm - t->t0) / t->dtt;
        W(s, 2) = clampf_pos(x, t->NP);
        W(s, 3) = clampf_pos(y, t->NT);
        const double ua = (b - a) / (grav * air->m_air_kg) * 6.02214076e23 / 1e4;   /* air molecules / cm^2 */
        const double yo = clampf_pos((t->o3lp0 - log(pm)) / t->o3dlp, t->NO3);
        const int io = (int)yo;
        const double xo3 = (t->o3[io] + (double)(t->o3[io + 1] - t->o3[io]) * (yo - (double)io)) * air->o3;
        const double xc = gx ? gx[(size_t)k * N + i] : co2, xn = gx ? gx[(size_t)(L + k) * N + i] : n2o;
        const double xm = gx ? gx[(size_t)(2 * L + k) * N + i] : ch4;
        W(s, 5) = (float)(xc * ua);
        W(s, 6) = (float)(xo3 * ua);
        W(s, 7) = (float)(xn * ua);
        W(s, 8) = (float)(xm * ua);
        W(s, 10) = 0.f;
        W(s, 11) = (float)fcl;
        W(s, 12) = (float)ua;
        W(s, 13) = (float)(pow(b / pb[L - 1], 8000.0 / hs) - pow(a / pb[L - 1], 8000.0 / hs));
        W(s, 14) = air->cia ? (float)(ua * 1e-20 * (pm / (1.380649e-23 * Tm) * 1e-6) * 1e-20) : 0.f;
        uas[j] = ua;
      }
#undef TOF
      /* the cloud in the sub-level holding the median of the layer's vapor */
      double cum = 0;
      int jc = 3;
      for (int j = 0; j < 4; j++) {
        cum += qm > 0 ? qs[j] * ((pe[j + 1] - pe[j]) / (pbk - pt)) / qm : 0.25;
        if (cum >= 0.5) {
          jc = j;
          break;
        }
      }
      if (bl_q > 0 && k == L - 1 && k > 0) jc = 0;
      W(k * 4 + jc, 10) = (float)tcl;
      for (int j = 0; j < 4; j++) {
        const int s = k * 4 + j;
        const double qj = qm > 0 ? q[k] * qs[j] / qm : q[k];
        const double x = qj / (1 - qj) * (air->m_air / 18.01528);   /* H2O volume mixing ratio */
        W(s, 4) = (float)(x * uas[j]);
        W(s, 9) = (float)(x / 0.04);
      }
    }
#undef W
  }
}

/* ---- lw_column (air.cu:2231-2348) ---- */

/* Through one sub-level of optical depth tau, from the far side's Planck radiance Bf to the near side's Bn, the
 * source linear in optical depth, one diffusivity angle (1/1.66) (air.cu:2123-2127). */
static double lw_through(double I, double tau, double Bn, double Bf) {
  const double x = tau * 1.66, tr = exp(-x);
  const double lin = x > 1e-3 ? 1 - (1 - tr) / x : x * (0.5 - x / 6);
  return I * tr + Bf * (1 - tr) + (Bn - Bf) * lin;
}

/* ln k at one table cell, bilinear in (log p, T), then exp (air.cu:2266-2271). */
static double kbil(const float* K, int NT, int NG, double fp, double ft) {
  const double v00 = K[0], v01 = K[NG], v10 = K[(size_t)NT * NG], v11 = K[(size_t)NT * NG + NG];
  return exp(v00 + (v10 - v00) * fp + (v01 - v00) * ft + (v00 - v01 - v10 + v11) * fp * ft);
}

void og_lw_column(int N, int L, int nsub, const double* par, const og_lwt* t, const float* sub, const double* Ts,
                  double* radf) {
  const int NG = t->NG, NP = t->NP, NT = t->NT, NTP = t->NTP, NS = nsub * L;
  const int RW = L + 1, E0 = 3 * RW + 3, rows = E0 + 6 * L;
  const double refl = par[P_LW_REFL] > 0 ? 1 - par[P_EPS_S] : 0.0;
  const size_t cs = (size_t)NP * NT * NG;
#pragma omp parallel
  {
    double *Tt = malloc(sizeof(double) * NS * 8), *Tb = Tt + NS, *fp = Tb + NS, *ft = fp + NS, *fh = ft + NS;
    double *tc = fh + NS, *fc = tc + NS, *wcia = fc + NS;
    double *u = malloc(sizeof(double) * NS * 5), *tau = malloc(sizeof(double) * NS);
    int *ip = malloc(sizeof(int) * NS * 2), *it = ip + NS;
    float* red = malloc(sizeof(float) * (size_t)rows * NG);
#pragma omp for schedule(static)
    for (int i = 0; i < N; i++) {
      for (int s = 0; s < NS; s++) {
#define R_(v) sub[((size_t)s * NV + (v)) * N + i]
        Tt[s] = R_(0);
        Tb[s] = R_(1);
        const double x = clampf_pos(R_(2), NP), y = clampf_pos(R_(3), NT);   /* (no-ops on lw_setup's positions) */
        ip[s] = (int)x;
        fp[s] = x - (double)ip[s];
        it[s] = (int)y;
        ft[s] = y - (double)it[s];
        for (int c = 0; c < 5; c++) u[s * 5 + c] = R_(4 + c);
        fh[s] = R_(9);
        tc[s] = R_(10);
        fc[s] = R_(11);
        wcia[s] = R_(14);
#undef R_
      }
      memset(red, 0, sizeof(float) * (size_t)rows * NG);
      for (int g = 0; g < NG; g++) {
#define RED(r) red[(size_t)(r) * NG + g]
        for (int s = 0; s < NS; s++) {
          const size_t b0 = ((size_t)ip[s] * NT + it[s]) * NG + g;
          double lk[6];
          for (int c = 0; c < 6; c++) lk[c] = kbil(t->logk + c * cs + b0, NT, NG, fp[s], ft[s]);
          const double* us = u + s * 5;
          tau[s] = fmax(0.0, us[0] * (lk[0] + (lk[1] - lk[0]) * fh[s]) + us[1] * lk[2] + us[2] * lk[3] + us[3] * lk[4] +
                                 us[4] * lk[5]);
          if (t->cia) {
            const float* C = t->cia + (size_t)it[s] * NG + g;
            tau[s] += wcia[s] * (C[0] + (C[NG] - C[0]) * ft[s]);
          }
        }
        /* Planck radiance per g-point, linear in T on the table (air.cu:2279-2284); the position rounds to float */
#define BOF(T, out)                                                                                           \
  do {                                                                                                        \
    float y_ = fminf(fmaxf((float)((T) - t->tp0), 0.f), (float)((double)(NTP - 1) - (double)1e-4f));          \
    const int j_ = (int)y_;                                                                                   \
    out = t->planck[(size_t)j_ * NG + g] +                                                                    \
          (double)(t->planck[(size_t)(j_ + 1) * NG + g] - t->planck[(size_t)j_ * NG + g]) * ((double)y_ - (double)j_); \
  } while (0)
        double I = 0, Icl = 0;
        RED(RW) = 0.f;
        for (int k = 0; k < L; k++) {
          const double f = fc[k * nsub];
          double Ic = I, Id = I;
          for (int s = k * nsub; s < (k + 1) * nsub; s++) {
            double Bt, Bb;
            BOF(Tt[s], Bt);
            BOF(Tb[s], Bb);
            Ic = lw_through(Ic, tau[s], Bb, Bt);
            if (refl > 0) Icl = lw_through(Icl, tau[s], Bb, Bt);
            if (f > 0) Id = lw_through(Id, tau[s] + tc[s] / 1.66, Bb, Bt);
            if (s == k * nsub) RED(E0 + 6 * k + 1) = (float)(PI_ * (f > 0 ? (1 - f) * Ic + f * Id : Ic));
            if (s == (k + 1) * nsub - 2) RED(E0 + 6 * k + 4) = (float)(PI_ * (f > 0 ? (1 - f) * Ic + f * Id : Ic));
          }
          I = f > 0 ? (1 - f) * Ic + f * Id : Ic;
          RED(RW + k + 1) = (float)(PI_ * I);
        }
        double Bs;
        BOF(Ts[i], Bs);
        double Iu = refl * I, tr = 1, Iuc = refl * Icl, trc = 1;
        RED(L) = (float)(PI_ * Iu);
        RED(2 * RW + L) = (float)Bs;
        for (int k = L - 1; k >= 0; k--) {
          const double f = fc[k * nsub];
          double Ic = Iu, Id = Iu, tcl = tr, tcd = tr;
          for (int s = (k + 1) * nsub - 1; s >= k * nsub; s--) {
            double Bt, Bb;
            BOF(Tt[s], Bt);
            BOF(Tb[s], Bb);
            const double t1 = exp(-tau[s] * 1.66);
            Iuc = lw_through(Iuc, tau[s], Bt, Bb);
            trc *= t1;
            Ic = lw_through(Ic, tau[s], Bt, Bb);
            tcl *= t1;
            if (f > 0) {
              Id = lw_through(Id, tau[s] + tc[s] / 1.66, Bt, Bb);
              tcd *= t1 * exp(-tc[s]);
            }
            const int e = s == (k + 1) * nsub - 1 ? 3 : (s == k * nsub + 1 ? 0 : -1);
            if (e >= 0) {
              RED(E0 + 6 * k + e) = (float)(PI_ * (f > 0 ? (1 - f) * Ic + f * Id : Ic));
              RED(E0 + 6 * k + e + 2) = (float)((f > 0 ? (1 - f) * tcl + f * tcd : tcl) * Bs);
            }
          }
          Iu = f > 0 ? (1 - f) * Ic + f * Id : Ic;
          tr = f > 0 ? (1 - f) * tcl + f * tcd : tcl;
          RED(k) = (float)(PI_ * Iu);
          RED(2 * RW + k) = (float)(tr * Bs);
        }
        RED(3 * RW) = (float)(PI_ * Iuc);
        RED(3 * RW + 1) = (float)(trc * Bs);
        RED(3 * RW + 2) = (float)Bs;
#undef BOF
#undef RED
      }
      /* sums over the g-points in order, in FP64; the surface's shares divided by its Planck sum */
      double Bsum = 0;
      for (int g = 0; g < NG; g++) Bsum += (double)red[(size_t)(3 * RW + 2) * NG + g];
      for (int r = 0; r < rows; r++) {
        double v = 0;
        for (int g = 0; g < NG; g++) v += (double)red[(size_t)r * NG + g];
        if ((r >= 2 * RW && r < 3 * RW) || r == 3 * RW + 1 || (r >= E0 && (r - E0) % 3 == 2)) v /= Bsum;
        radf[(size_t)r * N + i] = v;
      }
    }
    free(Tt);
    free(u);
    free(tau);
    free(ip);
    free(red);
  }
}

/* ---- sw_layer (air.cu:2358-2382): Delta-Eddington reflectances and transmittances of one homogeneous layer for
 * diffuse (Rd, Td) and direct (Rs, Ts, Tdd) light from the sun at cosine mu0 ---- */
static void sw_layer(double tau, double ssa, double g, double mu0, double* Rd, double* Td, double* Rs, double* Ts,
                     double* Tdd) {
  const double f = g * g;
  tau = (1 - ssa * f) * tau;
  ssa = (1 - f) * ssa / (1 - ssa * f);
  g = (g - f) / (1 - f);
  const double g1 = (7 - ssa * (4 + 3 * g)) / 4, g2 = -(1 - ssa * (4 - 3 * g)) / 4, g3 = (2 - 3 * g * mu0) / 4;
  const double g4 = 1 - g3;
  const double a1 = g1 * g4 + g2 * g3, a2 = g1 * g3 + g2 * g4;
  const double k = sqrt(fmax((g1 - g2) * (g1 + g2), 1e-12));
  const double e = exp(-k * tau), e2 = e * e, k2e = 2 * k * e;
  double rf = 1 / (k + g1 + (k - g1) * e2);
  *Rd = g2 * (1 - e2) * rf;
  *Td = k2e * rf;
  *Tdd = exp(-tau / mu0);
  const double km = k * mu0;
  double om = 1 - km * km;
  if (fabs(om) < 1e-3) om = om < 0 ? -1e-3 : 1e-3;
  rf = ssa * rf / om;
  double rs = rf * ((1 - km) * (a2 + k * g3) - (1 + km) * (a2 - k * g3) * e2 - k2e * (g3 - a2 * mu0) * *Tdd);
  double ts = rf * (k2e * (g4 + a1 * mu0) - *Tdd * ((1 + km) * (a1 + k * g4) - (1 - km) * (a1 - k * g4) * e2));
  rs = fmin(fmax(rs, 0.0), 1 - *Tdd);
  ts = fmin(fmax(ts, 0.0), 1 - *Tdd - rs);
  *Rs = rs;
  *Ts = ts;
}

void og_sw_layer(double tau, double ssa, double g, double mu0, double* out) {
  sw_layer(tau, ssa, g, mu0, out, out + 1, out + 2, out + 3, out + 4);
}

/* Adding (air.cu:2460-2485): the albedos of what lies below each level (Ad diffuse, As direct), then the fluxes down
 * from the top; absk per layer, et/eb its top and bottom sub-level's absorption, surf the surface's; returns As[0]. */
static double sw_add(int NS, int L, int nsub, double a, const double* Rd, const double* Td, const double* Rs,
                     const double* Ts, const double* Tdd, double* Ad, double* As, double* absk, double* et, double* eb,
                     double* surf) {
  Ad[NS] = a;
  As[NS] = a;
  for (int s = NS - 1; s >= 0; s--) {
    const double den = 1 / (1 - Ad[s + 1] * Rd[s]);
    Ad[s] = Rd[s] + Td[s] * Td[s] * Ad[s + 1] * den;
    As[s] = Rs[s] + (Tdd[s] * As[s + 1] + Ts[s] * Ad[s + 1]) * Td[s] * den;
  }
  double Fdir = 1, Fdif = 0, net = 1 - As[0];
  for (int k = 0; k < L; k++) absk[k] = 0;
  for (int s = 0; s < NS; s++) {
    const double den = 1 / (1 - Ad[s + 1] * Rd[s]);
    const double Fdn = Fdir * Tdd[s];
    Fdif = (Td[s] * Fdif + Rd[s] * As[s + 1] * Fdn + Ts[s] * Fdir) * den;
    Fdir = Fdn;
    const double n2 = Fdir + Fdif - (Ad[s + 1] * Fdif + As[s + 1] * Fdir);
    absk[s / nsub] += net - n2;
    if (s % nsub == 0) et[s / nsub] = net - n2;
    if (s % nsub == nsub - 1) eb[s / nsub] = net - n2;
    net = n2;
  }
  *surf = net;
  return As[0];
}

/* ---- sw_column (air.cu:2384-2511) ---- */
void og_sw_column(int N, int L, int nsub, const double* par, const og_swt* t, const og_air* air, const float* sub,
                  int mode, double tnow, double interval, const double* pos, const double* alb, const double* cloudo,
                  double* swf) {
  const int NG = t->NG, NP = t->NP, NT = t->NT, NS = nsub * L, rows = 3 * L + 2;
  const size_t cs = (size_t)NP * NT * NG;
  const double cloud_g = par[P_CLOUD_G], tilt = par[P_TILT], day_s = par[P_DAY];
#pragma omp parallel
  {
    double* w = malloc(sizeof(double) * ((size_t)NS * 26 + 2 * (NS + 1) + 8 * L + 8));
    double *fp = w, *ft = fp + NS, *fh = ft + NS, *sr = fh + NS, *sh = sr + NS, *tcs = sh + NS, *fcs = tcs + NS;
    double *sli = fcs + NS, *wcia = sli + NS, *u = wcia + NS;                    /* u: NS x 6 */
    double *Rd = u + 6 * NS, *Td = Rd + NS, *Rs = Td + NS, *Ts = Rs + NS, *Tdd = Ts + NS;
    double *cRd = Tdd + NS, *cTd = cRd + NS, *cRs = cTd + NS, *cTs = cRs + NS, *cTdd = cTs + NS;
    double *Ad = cTdd + NS, *As = Ad + NS + 1, *absk = As + NS + 1, *et = absk + L, *eb = et + L, *cabs = eb + L;
    double *cet = cabs + L, *ceb = cet + L;
    int* ip = malloc(sizeof(int) * NS * 2);
    int* it = ip + NS;
    float* red = malloc(sizeof(float) * (size_t)rows * NG);
#pragma omp for schedule(static)
    for (int i = 0; i < N; i++) {
      for (int s = 0; s < NS; s++) {
        const int k = s / nsub;
#define R_(v) sub[((size_t)s * NV + (v)) * N + i]
        const double x = clampf_pos(R_(2), NP);   /* (a no-op when the tables share their pressure axis) */
        /* fminf of a double product: rounded to float */
        const double y = fmaxf(fminf((float)((double)R_(3) * t->yscale), (float)((double)(NT - 1) - (double)1e-4f)), 0.f);
        ip[s] = (int)x;
        fp[s] = x - (double)ip[s];
        it[s] = (int)y;
        ft[s] = y - (double)it[s];
        for (int c = 0; c < 5; c++) u[s * 6 + c] = R_(4 + c);
        u[s * 6 + 5] = air->o2 * R_(12);
        fh[s] = R_(9);
        sr[s] = R_(12);
        sh[s] = R_(13);
        wcia[s] = R_(14);
        const int cl = R_(10) > 0 && cloudo;
        fcs[s] = cl ? cloudo[(size_t)k * N + i] : 0.0;
        tcs[s] = cl ? t->inhom * cloudo[(size_t)(2 * L + 3 + k) * N + i] : 0.0;
        sli[s] = cl && t->cssa_i ? cloudo[(size_t)(3 * L + 3 + k) * N + i] : 1.0;
#undef R_
      }
      /* the sun: the mean of its cosine over the interval's daylight, four samples */
      double m = 0.5;
      if (mode == 0) {
        double sum = 0;
        int lit = 0;
        for (int j = 0; j < 4; j++) {
          const double tt = tnow + (j + 0.5) * interval / 4;
          double fd;
          const double ang = og_sun_lon(tt, par, &fd), dec = asin(sin(tilt) * sin(ang));
          const double rot = 2 * PI_ * fmod(tt, day_s) / day_s;
          const double cz = pos[3 * i] * cos(dec) * cos(-rot) + pos[3 * i + 1] * cos(dec) * sin(-rot) +
                            pos[3 * i + 2] * sin(dec);
          if (cz > 0) {
            sum += cz;
            lit++;
          }
        }
        m = lit ? sum / lit : 0.05;
      }
      const double mu0 = fmax(m, 0.02), a = alb[i];
      memset(red, 0, sizeof(float) * (size_t)rows * NG);
      for (int g = 0; g < NG; g++) {
        for (int s = 0; s < NS; s++) {
          const size_t b0 = ((size_t)ip[s] * NT + it[s]) * NG + g;
          double lk[7];
          for (int c = 0; c < 7; c++) lk[c] = kbil(t->logk + c * cs + b0, NT, NG, fp[s], ft[s]);
          const double* us = u + s * 6;
          double tg = fmax(0.0, us[0] * (lk[0] + (lk[1] - lk[0]) * fh[s]) + us[1] * lk[2] + us[2] * lk[3] +
                                    us[3] * lk[4] + us[4] * lk[5] + us[5] * lk[6]);
          if (t->cia) {
            const float* C = t->cia + (size_t)it[s] * NG + g;
            tg += wcia[s] * (C[0] + (C[NG] - C[0]) * ft[s]);
          }
          const double tr = t->ray[g] * sr[s] * air->ray, th = t->tau550 * t->hzf[g] * sh[s];
          const double tt = tg + tr + th, sc = tr + t->hssa * th;
          const double ssa = tt > 0 ? fmin(sc / tt, 1 - 1e-6) : 0.0;
          const double ga = sc > 0 ? t->hg * t->hssa * th / sc : 0.0;
          sw_layer(tt, ssa, ga, mu0, &cRd[s], &cTd[s], &cRs[s], &cTs[s], &cTdd[s]);
          const double f = fcs[s];
          if (f > 0 && tcs[s] > 0) {                  /* the cloudy part: the cloud's scattering added */
            const double tc = tcs[s], t2 = tt + tc;
            double cw = t->cssa[g], cg = cloud_g;
            if (t->cssa_i) {
              const double sl = sli[s], wl = sl * t->cssa[g], wi = (1 - sl) * t->cssa_i[g];
              cw = wl + wi;
              cg = cw > 0 ? (wl * cloud_g + wi * t->g_i[g]) / cw : cg;
            }
            const double sc2 = sc + cw * tc;
            const double g2 = (ga * sc + cg * cw * tc) / sc2;
            double r1, t1, r2, t3, d1;
            sw_layer(t2, fmin(sc2 / t2, 1 - 1e-6), g2, mu0, &r1, &t1, &r2, &t3, &d1);
            Rd[s] = (1 - f) * cRd[s] + f * r1;
            Td[s] = (1 - f) * cTd[s] + f * t1;
            Rs[s] = (1 - f) * cRs[s] + f * r2;
            Ts[s] = (1 - f) * cTs[s] + f * t3;
            Tdd[s] = (1 - f) * cTdd[s] + f * d1;
          } else {
            Rd[s] = cRd[s];
            Td[s] = cTd[s];
            Rs[s] = cRs[s];
            Ts[s] = cTs[s];
            Tdd[s] = cTdd[s];
          }
        }
        double surf, csurf;
        sw_add(NS, L, nsub, a, Rd, Td, Rs, Ts, Tdd, Ad, As, absk, et, eb, &surf);
        const double crefl = sw_add(NS, L, nsub, a, cRd, cTd, cRs, cTs, cTdd, Ad, As, cabs, cet, ceb, &csurf);
        const double wg = t->sun[g];
        for (int k = 0; k < L; k++) {
          red[(size_t)k * NG + g] = (float)(wg * absk[k]);
          red[(size_t)(L + 2 + 2 * k) * NG + g] = (float)(wg * et[k]);
          red[(size_t)(L + 3 + 2 * k) * NG + g] = (float)(wg * eb[k]);
        }
        red[(size_t)L * NG + g] = (float)(wg * surf);
        red[(size_t)(L + 1) * NG + g] = (float)(wg * crefl);
      }
      for (int r = 0; r < rows; r++) {
        double v = 0;
        for (int g = 0; g < NG; g++) v += (double)red[(size_t)r * NG + g];
        swf[(size_t)r * N + i] = v;
      }
    }
    free(w);
    free(ip);
    free(red);
  }
}
