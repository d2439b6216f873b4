// The gravity waves' drag on the air (reference: kernels/drag.cu, orodrag.py:38-96): the mountains' drag of Lott &
// Miller (1997) from each cell's sub-grid orography, and the spectrum of gravity waves not tied to the ground of
// Alexander & Dunkerton (1999). FP64 (the reference's double path), compiled with -ffp-contract=off, every value in
// the reference's operation order.
//
// Design (MI355X): the cell passes are column passes -- a cell's whole layer stack on one thread, the layer count a
// template parameter so that every per-layer array has a compile-time size (registers), 64-thread blocks dealt to
// the 8 XCDs in contiguous ranges (641 waves at level 6); they are bound by their own dependent arithmetic (pow,
// log, atan2 and, for the waves, 167 phase speeds with an exp and a search over the layers each), not by HBM. The
// edge pass is one thread per edge over the layers: it reads both cells' forces and factors (a gather, 8-byte words)
// and updates the winds and the dissipation in place. The planet's mean convective rain for the waves' convective
// source is one 256-thread block summing in the reference's fixed order, so it repeats bit for bit.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "diag.hpp"
#include "insilica_hip.h"

namespace isw_drag {

constexpr int BS1 = 64;
constexpr int BS = 256;

__device__ __forceinline__ int xblock() {
  const int nb = (int)gridDim.x, b = (int)blockIdx.x;
  const int q = nb >> 3, r = nb & 7, x = b & 7, k = b >> 3;
  return (x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + k;
}

// The primitive F(t) of sqrt((A - t) / t) on [0, A], t clamped there (drag.cu:32-35).
__device__ __forceinline__ double blk_int(double t, double A) {
  t = fmin(fmax(t, 0.0), A);
  return sqrt(t * (A - t)) + A * asin(sqrt(t / A));
}

// Layer k's wind (east, north) at cell i from its edges' normal winds (Perot's coefficients op), to the first pad.
__device__ __forceinline__ void cell_wind(int E, int i, int k, const int* __restrict__ cellEdge,
                                          const double* __restrict__ op, const double* __restrict__ u, double& ue,
                                          double& un) {
  double a = 0, b = 0;
  for (int m = 0; m < 6; m++) {
    const int e = cellEdge[i * 6 + m];
    if (e < 0) break;
    const double w = u[(size_t)k * E + e];
    a += op[(size_t)(i * 6 + m) * 2] * w;
    b += op[(size_t)(i * 6 + m) * 2 + 1] * w;
  }
  ue = a;
  un = b;
}

// The mountains' drag per cell (drag.cu:40-170): each layer's implicit blocked-flow factor beta (dt x rate) and the
// gravity waves' force per area F (east, north rows per layer).
template <int L>
__global__ __launch_bounds__(BS1) void k_oro_cell(int N, int E, isw_oro P, const double* __restrict__ mu,
                                                  const double* __restrict__ theta, const double* __restrict__ u,
                                                  const int* __restrict__ cellEdge, const double* __restrict__ op,
                                                  const double* __restrict__ sso, double* __restrict__ beta,
                                                  double* __restrict__ F) {
  const int i = xblock() * BS1 + (int)threadIdx.x;
  if (i >= N) return;
  double bk[L], fx[L], fy[L];
#pragma unroll
  for (int k = 0; k < L; k++) bk[k] = fx[k] = fy[k] = 0;
  const double ms = sso[i], sig = sso[N + i], tho = sso[3 * N + i];
  const double gam = fmax(sso[2 * N + i], 1e-7);
  if (ms > P.mu_min && sig > 0) {
    double U[L], V[L], zb[L], zt[L], zm[L], rho[L], nn[L], ri[L], pt[L], pb[L];
#pragma unroll
    for (int k = 0; k < L; k++) cell_wind(E, i, k, cellEdge, op, u, U[k], V[k]);
    double p = 0;
#pragma unroll
    for (int k = 0; k < L; k++) {
      pt[k] = p;
      p += P.grav * mu[(size_t)k * N + i];
      pb[k] = p;
    }
    double z = 0;
#pragma unroll
    for (int k = L - 1; k >= 0; k--) {
      const double pm = 0.5 * (pt[k] + pb[k]);
      const double T = theta[(size_t)k * N + i] * pow(pm / P.p0, P.kappa);
      rho[k] = pm / (P.rd * T);
      const double Hs = P.rd * T / P.grav;
      zb[k] = z;
      zm[k] = z + Hs * log(pb[k] / pm);
      zt[k] = k > 0 ? z + Hs * log(pb[k] / fmax(pt[k], 1e-3 * pb[k])) : 1e30;
      z = zt[k];
    }
#pragma unroll
    for (int k = 0; k < L; k++) {
      double ns = 0, ss = 0;
      int c = 0;
#pragma unroll
      for (int j = k; j <= k + 1; j++) {
        if (j < 1 || j > L - 1) continue;
        const double dz = fmax(zm[j - 1] - zm[j], 1.0);
        const double t0 = theta[(size_t)(j - 1) * N + i], t1 = theta[(size_t)j * N + i];
        ns += P.grav * (t0 - t1) / (0.5 * (t0 + t1) * dz);
        const double du = U[j - 1] - U[j], dv = V[j - 1] - V[j];
        ss += (du * du + dv * dv) / (dz * dz);
        c++;
      }
      const double s2 = ss / c, n2 = fmax(ns / c, 1e-6);
      nn[k] = sqrt(n2);
      ri[k] = fmax(n2 / fmax(s2, 1e-12), P.ric);
    }
    const double H = 2 * ms;
    double wsum = 0, ul = 0, vl = 0, nl = 0, rl = 0;
#pragma unroll
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
    double sl = 0;
    if (wsum > 0) {
      ul /= wsum;
      vl /= wsum;
      nl /= wsum;
      rl /= wsum;
      sl = sqrt(ul * ul + vl * vl);
    }
    if (wsum > 0 && !(sl < 0.1)) {
      const double B = 1 - 0.18 * gam - 0.04 * gam * gam, C = 0.48 * gam + 0.3 * gam * gam;
      const double psi = tho - atan2(vl, ul), sp = sin(psi), cs = cos(psi);
      const double d1 = B - (B - C) * sp * sp, d2 = (B - C) * sp * cs, dm = sqrt(d1 * d1 + d2 * d2);
      const double ex = (ul * d1 - vl * d2) / (sl * dm), ey = (vl * d1 + ul * d2) / (sl * dm);
      double up[L];
#pragma unroll
      for (int k = 0; k < L; k++) up[k] = U[k] * ex + V[k] * ey;
      const double upl = ul * ex + vl * ey;
      // the blocked layer's top: N / U_p integrated down from H reaches H_nc
      double acc = 0, zblk = 0;
      bool found = false;
#pragma unroll
      for (int k = 0; k < L; k++) {
        const double top = fmin(zt[k], H), bot = zb[k];
        if (found || top <= bot) continue;
        if (up[k] <= 1e-3) {
          zblk = top;
          found = true;
          continue;
        }
        const double rate = nn[k] / up[k];
        if (acc + rate * (top - bot) >= P.hnc) {
          zblk = top - (P.hnc - acc) / rate;
          found = true;
          continue;
        }
        acc += rate * (top - bot);
      }
      zblk = fmin(fmax(zblk, 0.0), H);
      if (zblk > 0) {
        const double A = zblk + ms;
        bool above = false;
#pragma unroll
        for (int k = L - 1; k >= 0; k--) {
          if (above || zb[k] >= zblk) {
            above = true;
            continue;
          }
          const double s = sqrt(U[k] * U[k] + V[k] * V[k]);
          if (s <= 0) continue;
          const double pk = tho - atan2(V[k], U[k]), c2 = cos(pk) * cos(pk), s2 = 1 - c2;
          const double r = (c2 + gam * s2) / (gam * c2 + s2);
          const double fac = fmax(2 - 1 / r, 0.0) * (B * c2 + C * s2);
          const double I = blk_int(fmin(zt[k], zblk) + ms, A) - blk_int(zb[k] + ms, A);
          const double dz = (k > 0 ? zt[k] : zblk) - zb[k];
          bk[k] = P.dt * P.cd * fac * sig / (4 * ms) * s * I / fmax(dz, 1.0);
        }
      }
      if (upl > 0) {
        // the waves launched at H, deposited going up (Lindzen's saturation; a critical level ends them)
        const double heff = H - zblk, q = sig * dm / (4 * ms);
        double tau = P.G * rl * nl * upl * q * heff * heff;
        int kh = 0;
        bool got = false;
#pragma unroll
        for (int k = L - 1; k >= 0; k--)
          if (!got && zt[k] > H) {
            kh = k;
            got = true;
          }
        const double scale = sl / upl;
        double fk[L];
#pragma unroll
        for (int k = 0; k < L; k++) fk[k] = 0;
        int prev = -1;
        bool done = false;
#pragma unroll
        for (int k = L - 1; k >= 0; k--) {
          if (done || k > kh || !(tau > 0)) continue;
          const double muk = mu[(size_t)k * N + i];
          const double room = up[k] > 0 ? muk * up[k] / P.dt : 0.0;
          if (up[k] <= 0) {
            // the critical level lies below this layer: the wave ends in the layer under it
#pragma unroll
            for (int j = 0; j < L; j++)
              if (j == prev) fk[j] = fmin(fk[j] + tau * scale, mu[(size_t)j * N + i] * up[j] / P.dt);
            tau = 0;
            done = true;
            continue;
          }
          double out;
          if (k == 0) {
            out = 0;
          } else {
            const double norm = P.G * rho[k] * nn[k] * up[k] * q;
            const double dh2 = tau / norm, a = nn[k] * sqrt(dh2) / up[k], sq = sqrt(ri[k]);
            const double riw = ri[k] * (1 - a) / ((1 + a * sq) * (1 + a * sq));
            out = tau;
            if (riw < P.ric) {
              const double del = 4 / sq / P.ric + 1 / (P.ric * P.ric) + 4 / P.ric, bb = 1 / P.ric + 2 / sq;
              const double ac = 0.5 * (-bb + sqrt(del));
              out = fmin(tau, norm * (up[k] * ac / nn[k]) * (up[k] * ac / nn[k]));
            }
          }
          const double f = fmin((tau - out) * scale, room);
          fk[k] = f;
          tau = k == 0 ? 0.0 : tau - f / scale;
          prev = k;
        }
#pragma unroll
        for (int k = 0; k < L; k++) {
          fx[k] = -fk[k] * ex;
          fy[k] = -fk[k] * ey;
        }
      }
    }
  }
#pragma unroll
  for (int k = 0; k < L; k++) {
    beta[(size_t)k * N + i] = bk[k];
    F[(size_t)(2 * k) * N + i] = fx[k];
    F[(size_t)(2 * k + 1) * N + i] = fy[k];
  }
}

// The waves not tied to the ground per cell (drag.cu:205-282): the eastward force per area of each layer (F's east
// rows; the north rows 0). cmean: the planet's mean convective rain (read when conv is given and fc > 0).
template <int L>
__global__ __launch_bounds__(BS1) void k_gw_cell(int N, int E, isw_gw P, int nc, const double* __restrict__ mu,
                                                 const double* __restrict__ theta, const double* __restrict__ u,
                                                 const int* __restrict__ cellEdge, const double* __restrict__ op,
                                                 const double* __restrict__ conv, const double* __restrict__ cmean,
                                                 double* __restrict__ F) {
  const int i = xblock() * BS1 + (int)threadIdx.x;
  if (i >= N) return;
  const double LN2 = 0.69314718055994531;
  double U[L], pt[L], pb[L], T[L], nn[L], dep[L];
  double p = 0;
#pragma unroll
  for (int k = 0; k < L; k++) {
    double a = 0;
    for (int m = 0; m < 6; m++) {
      const int e = cellEdge[i * 6 + m];
      if (e < 0) break;
      a += op[(size_t)(i * 6 + m) * 2] * u[(size_t)k * E + e];
    }
    U[k] = a;
    pt[k] = p;
    p += P.grav * mu[(size_t)k * N + i];
    pb[k] = p;
    T[k] = theta[(size_t)k * N + i] * pow(0.5 * (pt[k] + pb[k]) / P.p0, P.kappa);
    dep[k] = 0;
  }
#pragma unroll
  for (int k = 0; k < L; k++) {
    double s = 0;
    int c = 0;
#pragma unroll
    for (int j = k; j <= k + 1; j++) {
      if (j < 1 || j > L - 1) continue;
      const double t0 = theta[(size_t)(j - 1) * N + i], t1 = theta[(size_t)j * N + i];
      const double pm0 = 0.5 * (pt[j - 1] + pb[j - 1]), pm1 = 0.5 * (pt[j] + pb[j]);
      const double dz = fmax(P.rd * 0.5 * (T[j - 1] + T[j]) / P.grav * log(pm1 / fmax(pm0, 1.0)), 1.0);
      s += P.grav * (t0 - t1) / (0.5 * (t0 + t1) * dz);
      c++;
    }
    nn[k] = sqrt(fmax(s / c, 1e-6));
  }
  int ks = -1;
#pragma unroll
  for (int k = L - 1; k >= 0; k--)
    if (pt[k] < P.psrc && P.psrc <= pb[k]) ks = k;       // (the first such layer from the top)
  if (ks > 0) {
    double us = 0, Tks = 0, nks = 0;
#pragma unroll
    for (int k = 0; k < L; k++)
      if (k == ks) {
        us = U[k];
        Tks = T[k];
        nks = nn[k];
      }
    const double rhos = P.psrc / (P.rd * Tks), Hs = P.rd * Tks / P.grav;
    const double wc = nks * P.kw / sqrt(P.kw * P.kw + 0.25 / (Hs * Hs));
    double thr[L];
#pragma unroll
    for (int k = 0; k < L; k++) {
      const double rho = fmax(pt[k], 1.0) / (P.rd * T[k]);
      thr[k] = 0.5 * (rho / rhos) * P.kw / nn[k];
    }
    double bsum = 0;
    for (int n = 0; n < nc; n++) {
      const double c = -P.cmax + n * P.dc, x = (c - us) / P.cw;
      if (c != us) bsum += exp(-LN2 * x * x);
    }
    if (bsum > 0) {
      const double amp =
          (conv != nullptr && P.fc > 0 ? P.bt * ((1 - P.fc) + P.fc * conv[i] / fmax(cmean[0], 1e-30)) : P.bt) / bsum;
      for (int n = 0; n < nc; n++) {
        const double c = -P.cmax + n * P.dc, cs = c - us;
        if (cs == 0 || fabs(cs) * P.kw >= wc) continue;
        const double x = cs / P.cw, sf = cs > 0 ? 1.0 : -1.0;
        const double b0 = sf * exp(-LN2 * x * x);
        // the layer the wave ends in: where it breaks or past its critical level, searching up from the source
        int kd = 0;
        bool found = false;
#pragma unroll
        for (int k = L - 1; k >= 1; k--) {
          if (found || k > ks) continue;
          const double cu = c - U[k];
          if (cs * cu <= 0) {
            kd = k + 1;
            found = true;
          } else if (b0 / (cu * cu * cu) >= thr[k]) {
            kd = k;
            found = true;
          }
        }
        if (kd == 0 && sf * (c - U[0]) <= 0) kd = 1;
#pragma unroll
        for (int k = 0; k < L; k++)
          if (k == kd) {
            const double room = fmax(0.0, sf * (c - U[k]) * mu[(size_t)k * N + i] / P.dt - sf * dep[k]);
            dep[k] += sf * fmin(fabs(amp * b0), room);
          }
      }
      double net = 0;
#pragma unroll
      for (int k = 0; k < L; k++) net += dep[k];
#pragma unroll
      for (int k = 0; k < L; k++)
        if (k == ks) dep[k] -= net;                      // (what no layer could take stays with the source)
    }
  }
#pragma unroll
  for (int k = 0; k < L; k++) {
    F[(size_t)(2 * k) * N + i] = dep[k];
    F[(size_t)(2 * k + 1) * N + i] = 0;
  }
}

// The planet's area-mean convective rain (drag.cu:285-293): 256 strided partial sums, then a halving tree.
__global__ __launch_bounds__(BS) void k_gw_mean(int N, const double* __restrict__ conv,
                                                const double* __restrict__ area, double* __restrict__ out) {
  __shared__ double sa[BS], sb[BS];
  const int t = (int)threadIdx.x;
  double a = 0, b = 0;
  for (int i = t; i < N; i += BS) {
    a += area[i] * conv[i];
    b += area[i];
  }
  sa[t] = a;
  sb[t] = b;
  __syncthreads();
  for (int s = BS / 2; s > 0; s >>= 1) {
    if (t < s) {
      sa[t] += sa[t + s];
      sb[t] += sb[t + s];
    }
    __syncthreads();
  }
  if (t == 0) out[0] = sa[0] / sb[0];
}

// Per edge and layer (drag.cu:172-193): the blocked drag (implicit, beta) and the force along the edge's normal
// (nrm: (east, north) at its first and second cell), the kinetic energy removed added to diss.
__global__ __launch_bounds__(BS) void k_drag_edge(int N, int E, int L, const double* __restrict__ mu,
                                                  const int* __restrict__ eI, const int* __restrict__ eJ,
                                                  const double* __restrict__ nrm, const double* __restrict__ beta,
                                                  const double* __restrict__ F, double dt, double* __restrict__ u,
                                                  double* __restrict__ diss) {
  const int e = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (e >= E) return;
  const int a = eI[e], b = eJ[e];
  const double4 n4 = reinterpret_cast<const double4*>(nrm)[e];
  double d = 0;
  for (int k = 0; k < L; k++) {
    const double m = 0.5 * (mu[(size_t)k * N + a] + mu[(size_t)k * N + b]);
    const double fn = 0.5 * (F[(size_t)(2 * k) * N + a] * n4.x + F[(size_t)(2 * k + 1) * N + a] * n4.y +
                             F[(size_t)(2 * k) * N + b] * n4.z + F[(size_t)(2 * k + 1) * N + b] * n4.w);
    const double bt = 0.5 * (beta[(size_t)k * N + a] + beta[(size_t)k * N + b]);
    if (fn == 0 && bt == 0) continue;
    double* pu = u + (size_t)k * E + e;
    const double u0 = *pu, u1 = (u0 + dt * fn / m) / (1 + bt);
    *pu = u1;
    d += 0.5 * m * (u0 * u0 - u1 * u1);
  }
  diss[e] += d;
}

}  // namespace isw_drag

using namespace isw_drag;

namespace {
bool drag_grid_ok(const isw_grid* g) {
  return g && g->N > 0 && g->E > 0 && g->eI && g->eJ && g->cellEdge && g->areaCell;
}
// phase speeds of the wave spectrum: -cmax .. cmax by dc (drag.cu:239)
int gw_count(const isw_gw* P) { return (int)std::lround(2 * P->cmax / P->dc) + 1; }
}  // namespace

extern "C" {

int32_t isw_oro_drag(const isw_grid* g, int32_t L, const isw_oro* par, const double* mu, const double* theta,
                     double* u, const double* op, const double* nrm, const double* sso, double* beta, double* F,
                     double dt, double* diss, void* stream) {
  if (!drag_grid_ok(g) || L < 2 || L > ISW_LMAX || !par || !mu || !theta || !u || !op || !nrm || !sso || !beta ||
      !F || !diss || !(dt > 0) || dt != par->dt || !(par->grav > 0) || !(par->p0 > 0) || !(par->rd > 0) ||
      !(par->ric > 0))
    return ISW_EBADARG;
  hipStream_t st = (hipStream_t)stream;
  const int N = g->N, E = g->E;
  const double n_ = N, e_ = E, l_ = L, D = 8, I = 4;
  const dim3 gc((unsigned)((N + BS1 - 1) / BS1));
#define ORO(LL)                                                                                                      \
  ISW_LAUNCH("oro_cell", (2 * l_ * n_ + l_ * e_ + 4 * n_ + 12 * n_ + 3 * l_ * n_) * D + 6 * n_ * I, k_oro_cell<LL>, \
             gc, dim3(BS1), N, E, *par, mu, theta, u, g->cellEdge, op, sso, beta, F)
  switch (L) {
    case 2: ORO(2); break;
    case 3: ORO(3); break;
    case 4: ORO(4); break;
    case 5: ORO(5); break;
    case 6: ORO(6); break;
    case 7: ORO(7); break;
    default: ORO(8); break;
  }
#undef ORO
  ISW_LAUNCH("drag_edge", (4 * l_ * n_ + 4 * e_ + 2 * l_ * e_ + 2 * e_) * D + 2 * e_ * I, k_drag_edge,
             dim3((unsigned)((E + BS - 1) / BS)), dim3(BS), N, E, (int)L, mu, g->eI, g->eJ, nrm, beta, F, dt, u, diss);
  return (int32_t)hipGetLastError();
}

int32_t isw_wave_drag(const isw_grid* g, int32_t L, const isw_gw* par, const double* mu, const double* theta,
                      double* u, const double* op, const double* nrm, const double* conv, double* cmean,
                      const double* zero, double* F, double dt, double* diss, void* stream) {
  if (!drag_grid_ok(g) || L < 2 || L > ISW_LMAX || !par || !mu || !theta || !u || !op || !nrm || !zero || !F ||
      !diss || !(dt > 0) || dt != par->dt || !(par->grav > 0) || !(par->p0 > 0) || !(par->rd > 0) ||
      !(par->cw > 0) || !(par->dc > 0) || !(par->cmax >= 0) || !(par->kw > 0) || (conv && !cmean))
    return ISW_EBADARG;
  const int nc = gw_count(par);
  if (nc < 1 || nc > ISW_GW_MAX_SPEEDS) return ISW_EBADARG;
  hipStream_t st = (hipStream_t)stream;
  const int N = g->N, E = g->E;
  const double n_ = N, e_ = E, l_ = L, D = 8, I = 4;
  if (conv) ISW_LAUNCH("gw_mean", 2 * n_ * D + D, k_gw_mean, dim3(1), dim3(BS), N, conv, g->areaCell, cmean);
  const dim3 gc((unsigned)((N + BS1 - 1) / BS1));
#define GW(LL)                                                                                                  \
  ISW_LAUNCH("gw_cell", (2 * l_ * n_ + l_ * e_ + 12 * n_ + 2 * l_ * n_ + (conv ? n_ : 0)) * D + 6 * n_ * I,    \
             k_gw_cell<LL>, gc, dim3(BS1), N, E, *par, nc, mu, theta, u, g->cellEdge, op, conv, cmean, F)
  switch (L) {
    case 2: GW(2); break;
    case 3: GW(3); break;
    case 4: GW(4); break;
    case 5: GW(5); break;
    case 6: GW(6); break;
    case 7: GW(7); break;
    default: GW(8); break;
  }
#undef GW
  ISW_LAUNCH("drag_edge", (4 * l_ * n_ + 4 * e_ + 2 * l_ * e_ + 2 * e_) * D + 2 * e_ * I, k_drag_edge,
             dim3((unsigned)((E + BS - 1) / BS)), dim3(BS), N, E, (int)L, mu, g->eI, g->eJ, nrm, zero, F, dt, u, diss);
  return (int32_t)hipGetLastError();
}

}  // extern "C"
