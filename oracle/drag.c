/* The CPU oracle of the gravity waves' drag on the air -- TEST INFRASTRUCTURE ONLY.
 *
 * A scalar C restatement of the reference's kernels/drag.cu (oro_cell :40-170, oro_edge :172-193, gw_cell :205-282,
 * gw_mean :285-293) as OroDrag.apply and WaveDrag.apply sequence them (orodrag.py:62-68, :86-96), in the reference's
 * operation order, FP64 (the reference's double path). OpenMP over cells and edges. The tests use it as the checker
 * of the HIP library's isw_oro_drag and isw_wave_drag.
 *
 * Layouts (the reference's): mu, theta, beta (L, N); u (L, E); F (2L, N), rows (east, north) per layer; op (N, 6, 2)
 * Perot's coefficients per edge slot; nrm (E, 4) the edge's normal (east, north) at its first and second cell;
 * sso (4, N): the sub-grid orography's standard deviation, slope, anisotropy and angle.
 */
#include <math.h>
#include <stddef.h>

#define OG_LMAX 8

/* oro par: grav, p0, kappa, rd, G, cd, hnc, ric, dt, mu_min */

/* The primitive F(t) of sqrt((A - t) / t) on [0, A] (t clamped there): sqrt(t (A - t)) + A asin(sqrt(t / A)). */
static double blk_int(double t, double A) {
  if (t < 0) t = 0;
  if (t > A) t = A;
  return sqrt(t * (A - t)) + A * asin(sqrt(t / A));
}

/* A cell's wind (east, north) in layer k from its edges' normal winds (Perot), stopping at the first pad. */
static void cell_wind(int N, int E, int k, int i, const int* cellEdge, const double* op, const double* u, double* ue,
                      double* un) {
  double a = 0, b = 0;
  for (int m = 0; m < 6; m++) {
    const int e = cellEdge[i * 6 + m];
    if (e < 0) break;
    const double w = u[(size_t)k * E + e];
    a += op[(size_t)(i * 6 + m) * 2] * w;
    b += op[(size_t)(i * 6 + m) * 2 + 1] * w;
  }
  (void)N;
  *ue = a;
  *un = b;
}

static void oro_one(int N, int E, int L, int i, const double* mu, const double* theta, const double* u,
                    const int* cellEdge, const double* op, const double* sso, const double* par, double* beta,
                    double* F) {
  const double grav = par[0], p0 = par[1], kappa = par[2], rd = par[3], Gw = par[4], cd = par[5], hnc = par[6],
               ric = par[7], dt = par[8], mu_min = par[9];
  for (int k = 0; k < L; k++) {
    beta[(size_t)k * N + i] = 0;
    F[(size_t)(2 * k) * N + i] = 0;
    F[(size_t)(2 * k + 1) * N + i] = 0;
  }
  const double ms = sso[i], sig = sso[N + i], tho = sso[3 * N + i];
  const double gam = fmax(sso[2 * N + i], 1e-7);
  if (!(ms > mu_min) || !(sig > 0)) return;
  double U[OG_LMAX], V[OG_LMAX], zb[OG_LMAX], zt[OG_LMAX], zm[OG_LMAX], rho[OG_LMAX], nn[OG_LMAX], ri[OG_LMAX];
  double pt[OG_LMAX], pb[OG_LMAX];
  for (int k = 0; k < L; k++) cell_wind(N, E, k, i, cellEdge, op, u, &U[k], &V[k]);
  double p = 0;
  for (int k = 0; k < L; k++) {
    pt[k] = p;
    p += grav * mu[(size_t)k * N + i];
    pb[k] = p;
  }
  /* heights above the ground, bottom layer first */
  double z = 0;
  for (int k = L - 1; k >= 0; k--) {
    const double pm = 0.5 * (pt[k] + pb[k]);
    const double T = theta[(size_t)k * N + i] * pow(pm / p0, kappa);
    rho[k] = pm / (rd * T);
    const double Hs = rd * T / grav;
    zb[k] = z;
    zm[k] = z + Hs * log(pb[k] / pm);
    zt[k] = k > 0 ? z + Hs * log(pb[k] / fmax(pt[k], 1e-3 * pb[k])) : 1e30;
    z = zt[k];
  }
  /* stability and shear across the interfaces above and below each layer */
  for (int k = 0; k < L; k++) {
    double ns = 0, ss = 0;
    int c = 0;
    for (int j = k; j <= k + 1; j++) {
      if (j < 1 || j > L - 1) continue;
      const double dz = fmax(zm[j - 1] - zm[j], 1.0);
      const double t0 = theta[(size_t)(j - 1) * N + i], t1 = theta[(size_t)j * N + i];
      ns += grav * (t0 - t1) / (0.5 * (t0 + t1) * dz);
      const double du = U[j - 1] - U[j], dv = V[j - 1] - V[j];
      ss += (du * du + dv * dv) / (dz * dz);
      c++;
    }
    const double s2 = ss / c, n2 = fmax(ns / c, 1e-6);
    nn[k] = sqrt(n2);
    ri[k] = fmax(n2 / fmax(s2, 1e-12), ric);
  }
  /* the low-level flow: mass-weighted means between mu and 2 mu */
  const double H = 2 * ms;
  double wsum = 0, ul = 0, vl = 0, nl = 0, rl = 0;
  for (int k = L - 1; k >= 0; k--) {
    const double ov = fmin(zt[k], H) - fmax(zb[k], ms);
    if (ov <= 0) continue;
    const double w = rho[k] * ov;
    wsum += w;
    ul += w * U[k];
    vl += w * V[k];
    nl += w * nn[k];
    rl += w * rho[k];
  }
  if (wsum <= 0) return;
  ul /= wsum;
  vl /= wsum;
  nl /= wsum;
  rl /= wsum;
  const double sl = sqrt(ul * ul + vl * vl);
  if (sl < 0.1) return;
  const double B = 1 - 0.18 * gam - 0.04 * gam * gam, C = 0.48 * gam + 0.3 * gam * gam;
  const double psi = tho - atan2(vl, ul), sp = sin(psi), cs = cos(psi);
  const double d1 = B - (B - C) * sp * sp, d2 = (B - C) * sp * cs, dm = sqrt(d1 * d1 + d2 * d2);
  const double ex = (ul * d1 - vl * d2) / (sl * dm), ey = (vl * d1 + ul * d2) / (sl * dm);
  double up[OG_LMAX];
  for (int k = 0; k < L; k++) up[k] = U[k] * ex + V[k] * ey;
  const double upl = ul * ex + vl * ey;
  /* the blocked layer's top */
  double acc = 0, zblk = 0;
  for (int k = 0; k < L; k++) {
    const double top = fmin(zt[k], H), bot = zb[k];
    if (top <= bot) continue;
    if (up[k] <= 1e-3) {
      zblk = top;
      break;
    }
    const double rate = nn[k] / up[k];
    if (acc + rate * (top - bot) >= hnc) {
      zblk = top - (hnc - acc) / rate;
      break;
    }
    acc += rate * (top - bot);
  }
  zblk = fmin(fmax(zblk, 0.0), H);
  if (zblk > 0) {
    const double A = zblk + ms;
    for (int k = L - 1; k >= 0; k--) {
      if (zb[k] >= zblk) break;
      const double s = sqrt(U[k] * U[k] + V[k] * V[k]);
      if (s <= 0) continue;
      const double pk = tho - atan2(V[k], U[k]), c2 = cos(pk) * cos(pk), s2 = 1 - c2;
      const double r = (c2 + gam * s2) / (gam * c2 + s2);
      const double fac = fmax(2 - 1 / r, 0.0) * (B * c2 + C * s2);
      const double I = blk_int(fmin(zt[k], zblk) + ms, A) - blk_int(zb[k] + ms, A);
      const double dz = (k > 0 ? zt[k] : zblk) - zb[k];
      beta[(size_t)k * N + i] = dt * cd * fac * sig / (4 * ms) * s * I / fmax(dz, 1.0);
    }
  }
  if (upl <= 0) return;
  const double heff = H - zblk, q = sig * dm / (4 * ms);
  double tau = Gw * rl * nl * upl * q * heff * heff;
  int kh = 0;
  for (int k = L - 1; k >= 0; k--)
    if (zt[k] > H) {
      kh = k;
      break;
    }
  const double scale = sl / upl;
  double fk[OG_LMAX];
  for (int k = 0; k < L; k++) fk[k] = 0;
  int prev = -1;
  for (int k = kh; k >= 0 && tau > 0; k--) {
    const double room = up[k] > 0 ? mu[(size_t)k * N + i] * up[k] / dt : 0.0;
    if (up[k] <= 0) {
      if (prev >= 0) fk[prev] = fmin(fk[prev] + tau * scale, mu[(size_t)prev * N + i] * up[prev] / dt);
      tau = 0;
      break;
    }
    double out;
    if (k == 0) {
      out = 0;
    } else {
      const double norm = Gw * rho[k] * nn[k] * up[k] * q;
      const double dh2 = tau / norm, a = nn[k] * sqrt(dh2) / up[k], sq = sqrt(ri[k]);
      const double riw = ri[k] * (1 - a) / ((1 + a * sq) * (1 + a * sq));
      out = tau;
      if (riw < ric) {
        const double del = 4 / sq / ric + 1 / (ric * ric) + 4 / ric, bb = 1 / ric + 2 / sq;
        const double ac = 0.5 * (-bb + sqrt(del));
        out = fmin(tau, norm * (up[k] * ac / nn[k]) * (up[k] * ac / nn[k]));
      }
    }
    const double f = fmin((tau - out) * scale, room);
    fk[k] = f;
    tau = k == 0 ? 0.0 : tau - f / scale;
    prev = k;
  }
  for (int k = 0; k < L; k++) {
    F[(size_t)(2 * k) * N + i] = -fk[k] * ex;
    F[(size_t)(2 * k + 1) * N + i] = -fk[k] * ey;
  }
}

/* The blocked drag (implicit) and the force F at each edge and layer; the kinetic energy removed is added to diss. */
static void drag_edges(int N, int E, int L, const double* mu, const int* eI, const int* eJ, const double* nrm,
                       const double* beta, const double* F, double dt, double* u, double* diss) {
#pragma omp parallel for schedule(static)
  for (int e = 0; e < E; e++) {
    const int a = eI[e], b = eJ[e];
    double d = 0;
    for (int k = 0; k < L; k++) {
      const double m = 0.5 * (mu[(size_t)k * N + a] + mu[(size_t)k * N + b]);
      const double fn = 0.5 * (F[(size_t)(2 * k) * N + a] * nrm[4 * (size_t)e] +
                               F[(size_t)(2 * k + 1) * N + a] * nrm[4 * (size_t)e + 1] +
                               F[(size_t)(2 * k) * N + b] * nrm[4 * (size_t)e + 2] +
                               F[(size_t)(2 * k + 1) * N + b] * nrm[4 * (size_t)e + 3]);
      const double bt = 0.5 * (beta[(size_t)k * N + a] + beta[(size_t)k * N + b]);
      if (fn == 0 && bt == 0) continue;
      double* pu = u + (size_t)k * E + e;
      const double u0 = *pu, u1 = (u0 + dt * fn / m) / (1 + bt);
      *pu = u1;
      d += 0.5 * m * (u0 * u0 - u1 * u1);
    }
    diss[e] += d;
  }
}

/* OroDrag.apply (orodrag.py:62-68): oro_cell then oro_edge. beta (L, N), F (2L, N): outputs (work). */
void og_oro_drag(int N, int E, int L, const int* cellEdge, const int* eI, const int* eJ, const double* mu,
                 const double* theta, double* u, const double* op, const double* nrm, const double* sso,
                 const double* par, double* beta, double* F, double dt, double* diss) {
#pragma omp parallel for schedule(static)
  for (int i = 0; i < N; i++) oro_one(N, E, L, i, mu, theta, u, cellEdge, op, sso, par, beta, F);
  drag_edges(N, E, L, mu, eI, eJ, nrm, beta, F, dt, u, diss);
}

/* gw par: grav, p0, kappa, rd, bt, cw, kw, psrc, cmax, dc, dt, fc */
static void gw_one(int N, int E, int L, int i, const double* mu, const double* theta, const double* u,
                   const int* cellEdge, const double* op, const double* par, const double* conv, double cmean,
                   double* F) {
  for (int k = 0; k < L; k++) {
    F[(size_t)(2 * k) * N + i] = 0;
    F[(size_t)(2 * k + 1) * N + i] = 0;
  }
  const double grav = par[0], p0 = par[1], kappa = par[2], rd = par[3], bt = par[4], cw = par[5], kw = par[6],
               psrc = par[7], cmax = par[8], dc = par[9], dt = par[10], fc = par[11];
  const double LN2 = 0.69314718055994531;
  double U[OG_LMAX], pt[OG_LMAX], pb[OG_LMAX], T[OG_LMAX], nn[OG_LMAX];
  double p = 0;
  for (int k = 0; k < L; k++) {
    double a = 0;
    for (int m = 0; m < 6; m++) {
      const int e = cellEdge[i * 6 + m];
      if (e < 0) break;
      a += op[(size_t)(i * 6 + m) * 2] * u[(size_t)k * E + e];
    }
    U[k] = a;
    pt[k] = p;
    p += grav * mu[(size_t)k * N + i];
    pb[k] = p;
    T[k] = theta[(size_t)k * N + i] * pow(0.5 * (pt[k] + pb[k]) / p0, kappa);
  }
  for (int k = 0; k < L; k++) {
    double s = 0;
    int c = 0;
    for (int j = k; j <= k + 1; j++) {
      if (j < 1 || j > L - 1) continue;
      const double t0 = theta[(size_t)(j - 1) * N + i], t1 = theta[(size_t)j * N + i];
      const double pm0 = 0.5 * (pt[j - 1] + pb[j - 1]), pm1 = 0.5 * (pt[j] + pb[j]);
      const double dz = fmax(rd * 0.5 * (T[j - 1] + T[j]) / grav * log(pm1 / fmax(pm0, 1.0)), 1.0);
      s += grav * (t0 - t1) / (0.5 * (t0 + t1) * dz);
      c++;
    }
    nn[k] = sqrt(fmax(s / c, 1e-6));
  }
  int ks = -1;
  for (int k = 0; k < L; k++)
    if (pt[k] < psrc && psrc <= pb[k]) {
      ks = k;
      break;
    }
  if (ks <= 0) return;
  const double us = U[ks], rhos = psrc / (rd * T[ks]), Hs = rd * T[ks] / grav;
  const double wc = nn[ks] * kw / sqrt(kw * kw + 0.25 / (Hs * Hs));
  const int nc = (int)round(2 * cmax / dc) + 1;
  double thr[OG_LMAX];
  for (int k = 0; k < L; k++) {
    const double rho = fmax(pt[k], 1.0) / (rd * T[k]);
    thr[k] = 0.5 * (rho / rhos) * kw / nn[k];
  }
  double bsum = 0;
  for (int n = 0; n < nc; n++) {
    const double c = -cmax + n * dc, x = (c - us) / cw;
    if (c != us) bsum += exp(-LN2 * x * x);
  }
  if (bsum <= 0) return;
  const double amp = (conv != NULL && fc > 0 ? bt * ((1 - fc) + fc * conv[i] / fmax(cmean, 1e-30)) : bt) / bsum;
  double dep[OG_LMAX];
  for (int k = 0; k < L; k++) dep[k] = 0;
  for (int n = 0; n < nc; n++) {
    const double c = -cmax + n * dc, cs = c - us;
    if (cs == 0 || fabs(cs) * kw >= wc) continue;
    const double x = cs / cw, sf = cs > 0 ? 1.0 : -1.0;
    const double b0 = sf * exp(-LN2 * x * x);
    int kd = 0;
    for (int k = ks; k >= 1; k--) {
      const double cu = c - U[k];
      if (cs * cu <= 0) {
        kd = k + 1;
        break;
      }
      if (b0 / (cu * cu * cu) >= thr[k]) {
        kd = k;
        break;
      }
    }
    if (kd == 0 && sf * (c - U[0]) <= 0) kd = 1;
    const double room = fmax(0.0, sf * (c - U[kd]) * mu[(size_t)kd * N + i] / dt - sf * dep[kd]);
    dep[kd] += sf * fmin(fabs(amp * b0), room);
  }
  double net = 0;
  for (int k = 0; k < L; k++) net += dep[k];
  dep[ks] -= net;
  for (int k = 0; k < L; k++) F[(size_t)(2 * k) * N + i] = dep[k];
}

/* gw_mean: the area mean of conv in the reference's order (256 strided partial sums, then a halving tree). */
double og_gw_mean(int N, const double* conv, const double* area) {
  double sa[256], sb[256];
  for (int t = 0; t < 256; t++) {
    double a = 0, b = 0;
    for (int i = t; i < N; i += 256) {
      a += area[i] * conv[i];
      b += area[i];
    }
    sa[t] = a;
    sb[t] = b;
  }
  for (int s = 128; s > 0; s >>= 1)
    for (int t = 0; t < s; t++) {
      sa[t] += sa[t + s];
      sb[t] += sb[t + s];
    }
  return sa[0] / sb[0];
}

/* WaveDrag.apply (orodrag.py:86-96): gw_mean (conv non-NULL), gw_cell, then the edges with no blocking. zero (L, N)
 * zeros; F (2L, N) output (work). */
void og_wave_drag(int N, int E, int L, const int* cellEdge, const int* eI, const int* eJ, const double* area,
                  const double* mu, const double* theta, double* u, const double* op, const double* nrm,
                  const double* par, const double* conv, const double* zero, double* F, double dt, double* diss) {
  const double cmean = conv ? og_gw_mean(N, conv, area) : 0.0;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < N; i++) gw_one(N, E, L, i, mu, theta, u, cellEdge, op, par, conv, cmean, F);
  drag_edges(N, E, L, mu, eI, eJ, nrm, zero, F, dt, u, diss);
}

/* The blocked drag's height integral, exported for the tests' quadrature pin. */
double og_blk_int(double t, double A) { return blk_int(t, A); }
