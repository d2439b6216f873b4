// The ocean's column processes as gfx950 HIP kernels: the reference's daily mixing of the mixed layer and the
// per-step exchanges with the air (src/insilica/kernels/ocean.cu, driven by ocean.py:226-295 Ocean.accumulate and
// Ocean.step(mix=True)). FP64 in both forms: these are N-sized passes, one thread per column, once per air step
// (sunlight, fresh water) or once per mixing interval (a model day), so they cost microseconds against the
// dynamics; compiled with -ffp-contract=off, every sum in the reference's order, so they match the reference's
// double path value for value.
//
// Kernels are instantiated per layer count L (2..8, ISW_LMAX) so that the per-layer arrays of a column (the
// adjustment's blocks, the interfaces' transfers) live in registers.
#include <hip/hip_runtime.h>

#include <cmath>

#include "diag.hpp"
#include "insilica_hip.h"

namespace isw_ocn {

constexpr int BS = 64;

__device__ __forceinline__ int xblock() {
  const int nb = (int)gridDim.x, b = (int)blockIdx.x;
  const int q = nb >> 3, r = nb & 7, x = b & 7, k = b >> 3;
  return (x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + k;
}
#define XTID ((int)(xblock() * BS + threadIdx.x))
static inline dim3 grid1(long n) { return dim3((unsigned)((n + BS - 1) / BS)); }

// The column physics' equation of state (ocean.cu:43-51): b = g (alpha th + alpha2 th^2 / 2 - beta S), th = T - 0 C.
struct Eos {
  double g, a, a2, beta;
};
__device__ __forceinline__ double buoy(double T, double S, const Eos& q) {
  const double th = T - 273.15;
  return q.g * (q.a * th + 0.5 * q.a2 * th * th - q.beta * S);
}
__device__ __forceinline__ double expand(double T, const Eos& q) { return q.a + q.a2 * (T - 273.15); }

// Convective adjustment of one column held in registers (ocean.cu:55-72): top down, each layer starts a block; while
// the block above is denser than the newest, the two merge (heat and salt mixed by thickness). Blocks live in a
// stack of at most L entries; a block's first layer marks where its values go back. Every stack access is written
// as a select over the (unrolled) slots, so the arrays stay in registers with compile-time indices only.
template <int L>
__device__ void adjust(const double (&h)[L], double (&T)[L], double (&S)[L], const Eos& q) {
  double hb[L], Tb[L], Sb[L];
  int first[L];
  int nb = 0;
#pragma unroll
  for (int k = 0; k < L; k++) {
#pragma unroll
    for (int j = 0; j < L; j++)
      if (j == nb) {
        hb[j] = h[k];
        Tb[j] = T[k];
        Sb[j] = S[k];
        first[j] = k;
      }
    nb++;
#pragma unroll
    for (int it = 0; it < k; it++) {               // at most k merges after the k-th push
      if (nb < 2) break;
      double ha = 0, Ta = 0, Sa = 0, hz = 0, Tz = 0, Sz = 0;
#pragma unroll
      for (int j = 0; j < L; j++) {
        if (j == nb - 2) { ha = hb[j]; Ta = Tb[j]; Sa = Sb[j]; }
        if (j == nb - 1) { hz = hb[j]; Tz = Tb[j]; Sz = Sb[j]; }
      }
      if (!(buoy(Ta, Sa, q) < buoy(Tz, Sz, q))) break;
      const double hs = ha + hz;
      const double Tn = (ha * Ta + hz * Tz) / hs, Sn = (ha * Sa + hz * Sz) / hs;
#pragma unroll
      for (int j = 0; j < L; j++)
        if (j == nb - 2) {
          Tb[j] = Tn;
          Sb[j] = Sn;
          hb[j] = hs;
        }
      nb--;
    }
  }
  // each layer takes its block's values: the block of layer k is the last whose first layer is <= k
#pragma unroll
  for (int k = 0; k < L; k++) {
    double t = Tb[0], s = Sb[0];
#pragma unroll
    for (int j = 1; j < L; j++)
      if (j < nb && first[j] <= k) {
        t = Tb[j];
        s = Sb[j];
      }
    T[k] = t;
    S[k] = s;
  }
}

// Paulson & Simpson's clear-water light passing depth z, and its integral from the surface (ocean.cu:75-79).
__device__ __forceinline__ double sun_f(double z, double R, double z1, double z2) {
  return R * exp(-z / z1) + (1 - R) * exp(-z / z2);
}
__device__ __forceinline__ double sun_int(double z, double R, double z1, double z2) {
  return -(R * z1 * expm1(-z / z1) + (1 - R) * z2 * expm1(-z / z2));
}
// The buoyancy flux a mixed layer of depth h works against, the sunlight counted where it is absorbed (:84-86).
__device__ __forceinline__ double kt_buoy(double h, double Bns, double Bsw, double R, double z1, double z2) {
  return Bns + Bsw * (1 + sun_f(h, R, z1, z2) - 2 * sun_int(h, R, z1, z2) / h);
}
// Kraus-Turner's turbulent energy at the base of a layer of depth h (:90-92).
__device__ __forceinline__ double kt_energy(double h, double us3, double B, double m_kt, double n_kt) {
  return 2 * m_kt * us3 - 0.5 * h * ((1 + n_kt) * B + (1 - n_kt) * fabs(B));
}

template <int L>
__global__ void __launch_bounds__(BS) k_ocn_overturn(int N, const double* __restrict__ h, double* __restrict__ T,
                                                     double* __restrict__ S, Eos q, const double* __restrict__ land) {
  const int i = XTID;
  if (i >= N) return;
  if (land && land[i] > 0.5) return;
  double hh[L], TT[L], SS[L];
#pragma unroll
  for (int k = 0; k < L; k++) {
    hh[k] = h[(size_t)k * N + i];
    TT[k] = T[(size_t)k * N + i];
    SS[k] = S[(size_t)k * N + i];
  }
  adjust<L>(hh, TT, SS, q);
#pragma unroll
  for (int k = 0; k < L; k++) {
    T[(size_t)k * N + i] = TT[k];
    S[(size_t)k * N + i] = SS[k];
  }
}

// The mixed layer over one mixing interval (ocn_column, ocean.cu:117-213): adjustment; Kraus-Turner entrainment at
// P / (h0 db) dt when the energy at the base P > 0 over a stratified base, else shoaling to the depth where P = 0
// (bisection); the base's shear instability (ri_b); upwelling to h_ent; the move bounded by the room above hmin (the
// layers below passed through with reach) and hmax; the interior interfaces' deep-water formation (dwf) or mixing
// (kappa_d); the water left behind joining the class of its own density (det_class). wx (L - 1, N), m, > 0 up.
struct MixPar {
  double S_ref, alpha_ref, rho_w, cw, m_kt, n_kt, hmin, hmax, dTmin, R, z1, z2, dt, kappa_d, dwf, ri_b, reach, h_ent,
      det_class;
};

template <int L>
__global__ void __launch_bounds__(BS) k_ocn_column(int N, const double* __restrict__ h, double* __restrict__ T,
                                                   double* __restrict__ S, const double* __restrict__ taum,
                                                   const double* __restrict__ qns, const double* __restrict__ qfw,
                                                   const double* __restrict__ qsun, Eos q, MixPar m,
                                                   double* __restrict__ wx, const double* __restrict__ land,
                                                   const double* __restrict__ shear, const double* __restrict__ Tc,
                                                   const double* __restrict__ Sc) {
  const int i = XTID;
  if (i >= N) return;
  if (land && land[i] > 0.5) {
#pragma unroll
    for (int k = 0; k < L - 1; k++) wx[(size_t)k * N + i] = 0;
    return;
  }
  double hh[L], TT[L], SS[L];
#pragma unroll
  for (int k = 0; k < L; k++) {
    hh[k] = h[(size_t)k * N + i];
    TT[k] = T[(size_t)k * N + i];
    SS[k] = S[(size_t)k * N + i];
  }
  adjust<L>(hh, TT, SS, q);
#pragma unroll
  for (int k = 0; k < L; k++) {
    T[(size_t)k * N + i] = TT[k];
    S[(size_t)k * N + i] = SS[k];
  }
  const double h0 = hh[0], h1 = hh[1];
  int kb = 1;                                      // (reach: the first layer below with water above the floor)
  if (m.reach > 0) {
    bool go = true;
#pragma unroll
    for (int k = 1; k < L - 1; k++) {
      go = go && hh[k] - m.hmin < 1;
      if (go) kb = k + 1;
    }
  }
  double T1 = TT[1], S1 = SS[1];
#pragma unroll
  for (int k = 2; k < L; k++)
    if (k == kb) {
      T1 = TT[k];
      S1 = SS[k];
    }
  const double T0 = TT[0];
  const double us = sqrt(taum[i] / m.rho_w);
  const double us3 = us * us * us;
  const double ga = q.g * expand(T0, q);
  const double Bns = ga * qns[i] / (m.rho_w * m.cw) + q.g * q.beta * m.S_ref * qfw[i] / m.rho_w;
  const double Bsw = ga * qsun[i] / (m.rho_w * m.cw);
  const double P = kt_energy(h0, us3, kt_buoy(h0, Bns, Bsw, m.R, m.z1, m.z2), m.m_kt, m.n_kt);
  const double db = buoy(T0, SS[0], q) - buoy(T1, S1, q);
  double dm = 0;
  if (P > 0) {
    if (db > q.g * m.alpha_ref * m.dTmin) dm = P / (h0 * db) * m.dt;
  } else {
    double lo = m.hmin, hi = h0;
    if (kt_energy(lo, us3, kt_buoy(lo, Bns, Bsw, m.R, m.z1, m.z2), m.m_kt, m.n_kt) > 0)
      for (int it = 0; it < 50; it++) {
        const double mid = 0.5 * (lo + hi);
        if (kt_energy(mid, us3, kt_buoy(mid, Bns, Bsw, m.R, m.z1, m.z2), m.m_kt, m.n_kt) > 0)
          lo = mid;
        else
          hi = mid;
      }
    dm = lo - h0;
  }
  if (shear && m.ri_b > 0 && db > 0) {
    const double du = shear[i];
    const double hn = sqrt(m.ri_b * du * du * h0 / db);
    if (hn - h0 > dm) dm = hn - h0;
  }
  if (m.h_ent > 0 && h0 + dm < m.h_ent) dm = m.h_ent - h0;
  double up_max = h1 - m.hmin;
  if (m.reach > 0) {
    up_max = 0;
#pragma unroll
    for (int k = 1; k < L; k++)
      if (k <= kb) up_max += fmax(hh[k] - m.hmin, 0.0);
  }
  if (m.hmax - h0 < up_max) up_max = m.hmax - h0;
  const double dn_max = h0 - m.hmin;
  if (dm > up_max) dm = up_max > 0 ? up_max : 0.0;
  if (-dm > dn_max) dm = dn_max > 0 ? -dn_max : 0.0;
  double w[L];                                     // the interfaces' transfers (w[k]: below layer k)
  w[0] = dm;
  double add[L + 1];                               // water passing up across the interface below layer k (reach)
#pragma unroll
  for (int k = 0; k <= L; k++) add[k] = 0;
  if (m.reach > 0 && dm > 0) {
    double rem = dm - fmax(h1 - m.hmin, 0.0);
#pragma unroll
    for (int k = 1; k < L - 1; k++)
      if (rem > 0) {
        add[k] = rem;
        rem -= fmax(hh[k + 1] - m.hmin, 0.0);
      }
  }
  double gave = dm > 0 ? dm - add[1] : 0.0;
#pragma unroll
  for (int k = 1; k < L - 1; k++) {
    const double hk = hh[k], hb = hh[k + 1];
    double v = 0;
    if (m.dwf > 0 && Bns < 0 && add[k] <= 0 && buoy(TT[k], SS[k], q) <= buoy(TT[k + 1], SS[k + 1], q))
      v = -fmax(hk - m.hmin - gave, 0.0);
    else if (m.kappa_d > 0)
      v = fmin(2 * m.kappa_d / (hk + hb) * m.dt, fmax(hb - m.hmin - (add[k] - add[k + 1]), 0.0));
    w[k] = v + add[k];
    gave = (v > 0 ? v : 0.0) + (add[k] - add[k + 1]);
  }
  if (Tc && m.det_class > 0 && dm < 0) {
    int ks = 1;
    const double b0 = buoy(T0, SS[0], q);
    for (int k = 2; k < L; k++)
      if (b0 <= 0.5 * (buoy(Tc[k - 1], Sc ? Sc[k - 1] : m.S_ref, q) + buoy(Tc[k], Sc ? Sc[k] : m.S_ref, q))) ks = k;
#pragma unroll
    for (int k = 1; k < L - 1; k++)
      if (k < ks) w[k] += dm;
  }
#pragma unroll
  for (int k = 0; k < L - 1; k++) wx[(size_t)k * N + i] = w[k];
}

// Salt moved with the water wx moves between layers (ocn_xch_salt, ocean.cu:352-381): upwind, conserved; seq: upward
// moves from the bottom interface up, downward from the top down, each at its donor's salinity as it then stands.
template <int L>
__global__ void __launch_bounds__(BS) k_ocn_xch_salt(int N, const double* __restrict__ wx,
                                                     const double* __restrict__ h, double* __restrict__ S, int seq) {
  const int i = XTID;
  if (i >= N) return;
  double hs[L], tr[L], s0[L], v[L];
#pragma unroll
  for (int k = 0; k < L; k++) {
    hs[k] = h[(size_t)k * N + i];
    s0[k] = S[(size_t)k * N + i];
    tr[k] = hs[k] * s0[k];
  }
#pragma unroll
  for (int k = 0; k < L - 1; k++) v[k] = wx[(size_t)k * N + i];
  if (seq) {
#pragma unroll
    for (int pass = 0; pass < 2; pass++)
#pragma unroll
      for (int n = 0; n + 1 < L; n++) {
        const int k = pass == 0 ? L - 2 - n : n;
        const double vk = v[k];
        if (pass == 0 ? !(vk > 0) : !(vk < 0)) continue;
        const double s = vk > 0 ? tr[k + 1] / hs[k + 1] : tr[k] / hs[k];
        tr[k] += vk * s;
        tr[k + 1] -= vk * s;
        hs[k] += vk;
        hs[k + 1] -= vk;
      }
  } else {
#pragma unroll
    for (int k = 0; k + 1 < L; k++) {
      const double s = v[k] > 0 ? s0[k + 1] : s0[k];
      tr[k] += v[k] * s;
      tr[k + 1] -= v[k] * s;
      hs[k] += v[k];
      hs[k + 1] -= v[k];
    }
  }
#pragma unroll
  for (int k = 0; k < L; k++) S[(size_t)k * N + i] = tr[k] / hs[k];
}

// The sunlight passing the mixed layer's base heats the layers below it (ocn_pen, ocean.cu:387-407), and the
// mixing's sums of the mixed layer's non-solar flux and sunlight; up (nullable): where water rose into the mixed
// layer at the last mixing it keeps all the sunlight.
template <int L>
__global__ void __launch_bounds__(BS) k_ocn_pen(int N, const double* __restrict__ h, double* __restrict__ T,
                                                const double* __restrict__ qocean, const double* __restrict__ qsw,
                                                double rho_cw, double R, double z1, double z2, double dt,
                                                double* __restrict__ qns_acc, double* __restrict__ sun_acc,
                                                const double* __restrict__ land, const double* __restrict__ up) {
  const int i = XTID;
  if (i >= N) return;
  if (land && land[i] > 0.5) return;
  const double qs = qsw[i];
  qns_acc[i] += (qocean[i] - qs) * dt;
  sun_acc[i] += qs * dt;
  if (up && up[i] > 0) return;
  double z = h[i];
  double f = sun_f(z, R, z1, z2);
  const double pen = qs * f;
  T[i] -= pen * dt / (rho_cw * h[i]);
#pragma unroll
  for (int k = 1; k < L; k++) {
    const double hk = h[(size_t)k * N + i];
    z += hk;
    const double fb = k < L - 1 ? sun_f(z, R, z1, z2) : 0.0;
    T[(size_t)k * N + i] += qs * (f - fb) * dt / (rho_cw * hk);
    f = fb;
  }
}

// Fresh water into the mixed layer every air step (ocn_fresh, ocean.cu:245-261): rain less evaporation less the water
// frozen into sea ice (plus what waits at a river mouth, emptied), as a virtual salt flux at S_ref, or (fw_local)
// with its volume, the salt kept; fw_acc sums it.
__global__ void __launch_bounds__(BS) k_ocn_fresh(int N, const double* __restrict__ h, double* __restrict__ S,
                                                  const double* __restrict__ evap, const double* __restrict__ rain,
                                                  const double* __restrict__ ice, double* __restrict__ ice_prev,
                                                  double S_ref, double rho_w, double dt, double* __restrict__ fw_acc,
                                                  const double* __restrict__ land, double* __restrict__ mouth,
                                                  double fw_local, double* __restrict__ hw, double* __restrict__ M,
                                                  const double* __restrict__ area) {
  const int i = XTID;
  if (i >= N) return;
  if (land && land[i] > 0.5) return;
  double fw = (rain[i] - evap[i]) * dt - (ice[i] - ice_prev[i]);
  if (mouth) {
    fw += mouth[i];
    mouth[i] = 0;
  }
  const double hi = h[i];
  if (fw_local > 0) {
    const double h1 = fmax(hi + fw / rho_w, fmin(hi, 10.0));
    S[i] *= hi / h1;
    hw[i] = h1;
    M[i] = h1 * area[i];
  } else {
    S[i] -= S_ref * fw / (rho_w * hi);
  }
  ice_prev[i] = ice[i];
  fw_acc[i] += fw;
}

// Mixing of T and S along each layer over the interval (ocn_diffuse, ocean.cu:268-292): across each open edge the flux
// kappa_h h_e (X_j - X_i) l / d, h_e the thinner side's thickness; (64, L) blocks over cells x layers.
__global__ void __launch_bounds__(BS) k_ocn_diffuse(int N, const double* __restrict__ h,
                                                    const double* __restrict__ T, const double* __restrict__ S,
                                                    double* __restrict__ Tn, double* __restrict__ Sn,
                                                    const int* __restrict__ cellEdge, const int* __restrict__ eI,
                                                    const int* __restrict__ eJ, const double* __restrict__ le,
                                                    const double* __restrict__ de,
                                                    const double* __restrict__ areaCell,
                                                    const double* __restrict__ wall, double kappa_h, double dt) {
  const int i = XTID;
  if (i >= N) return;
  const size_t a = (size_t)blockIdx.y * N;
  const double hi = h[a + i], Ti = T[a + i], Si = S[a + i];
  double fT = 0, fS = 0;
  for (int m = 0; m < 6; m++) {
    const int e = cellEdge[i * 6 + m];
    if (e < 0) break;
    if (wall && wall[e] == 0) continue;
    const int j = eI[e] == i ? eJ[e] : eI[e];
    const double he = fmin(hi, h[a + j]), c = kappa_h * he * le[e] / de[e] * dt;
    fT += c * (T[a + j] - Ti);
    fS += c * (S[a + j] - Si);
  }
  Tn[a + i] = Ti + fT / (hi * areaCell[i]);
  Sn[a + i] = Si + fS / (hi * areaCell[i]);
}

// The mixing interval's forcing (Ocean.step, ocean.py:257-266): the edges' rms stress sqrt(tau_acc (1 / dt)) in place,
// its cell speed into taum (as cell_speed forms a speed from edge values, air.cu:1978-1989); and the mean fluxes
// q_acc / dt, sun_acc / dt, fw_acc / dt in place.
__global__ void __launch_bounds__(BS) k_ocn_rms(int E, double rdt, double* __restrict__ tau_acc) {
  const int e = XTID;
  if (e >= E) return;
  tau_acc[e] = sqrt(tau_acc[e] * rdt);
}
__global__ void __launch_bounds__(BS) k_ocn_means(int N, const double* __restrict__ tau, const int* __restrict__ cellEdge,
                                                  const double* __restrict__ le, const double* __restrict__ de,
                                                  const double* __restrict__ areaCell, double* __restrict__ taum,
                                                  double dt, double* __restrict__ q_acc, double* __restrict__ sun_acc,
                                                  double* __restrict__ fw_acc) {
  const int i = XTID;
  if (i >= N) return;
  double ke = 0;
  for (int m = 0; m < 6; m++) {
    const int e = cellEdge[i * 6 + m];
    if (e < 0) break;
    ke += le[e] * de[e] * 0.25 * tau[e] * tau[e];
  }
  taum[i] = sqrt(2 * ke / areaCell[i]);
  q_acc[i] /= dt;
  sun_acc[i] /= dt;
  fw_acc[i] /= dt;
}

// Host side: dispatch on L.
#define ISW_BY_L(LV, CALL)            \
  switch (LV) {                       \
    case 2: { constexpr int LL = 2; CALL; } break; \
    case 3: { constexpr int LL = 3; CALL; } break; \
    case 4: { constexpr int LL = 4; CALL; } break; \
    case 5: { constexpr int LL = 5; CALL; } break; \
    case 6: { constexpr int LL = 6; CALL; } break; \
    case 7: { constexpr int LL = 7; CALL; } break; \
    case 8: { constexpr int LL = 8; CALL; } break; \
    default: return ISW_EBADARG;      \
  }

static Eos eos_of(const isw_ocean* o) { return Eos{o->grav, o->alpha, o->alpha2, o->beta}; }

int overturn(int N, int L, const isw_ocean* o, const double* h, double* T, double* S, const double* land,
             hipStream_t st) {
  const Eos q = eos_of(o);
  const double bytes = (double)N * 8 * (L + 4 * L + 1);
  ISW_BY_L(L, ISW_LAUNCH("ocn_overturn", bytes, k_ocn_overturn<LL>, grid1(N), dim3(BS), N, h, T, S, q, land));
  return (int)hipGetLastError();
}

int column(int N, int L, const isw_ocean* o, const double* h, double* T, double* S, const double* taum,
           const double* qns, const double* qfw, const double* qsun, double dt, double* wx, const double* land,
           const double* shear, const double* Tc, const double* Sc, hipStream_t st) {
  const Eos q = eos_of(o);
  const MixPar m{o->S_ref, o->alpha_ref, o->rho_w, o->cw, o->m_kt, o->n_kt, o->hmin, o->hmax, o->dTmin, o->sw_R,
                 o->sw_z1, o->sw_z2, dt, o->kappa_d, o->dwf, o->ri_b, o->reach, o->h_ent, o->det_class};
  const double bytes = (double)N * 8 * (L + 4 * L + 5 + (L - 1) + (shear ? 1 : 0));
  ISW_BY_L(L, ISW_LAUNCH("ocn_column", bytes, k_ocn_column<LL>, grid1(N), dim3(BS), N, h, T, S, taum, qns, qfw, qsun,
                         q, m, wx, land, shear, Tc, Sc));
  return (int)hipGetLastError();
}

int xch_salt(int N, int L, const double* wx, const double* h, double* S, int seq, hipStream_t st) {
  const double bytes = (double)N * 8 * ((L - 1) + L + 2 * L);
  ISW_BY_L(L, ISW_LAUNCH("ocn_xch_salt", bytes, k_ocn_xch_salt<LL>, grid1(N), dim3(BS), N, wx, h, S, seq));
  return (int)hipGetLastError();
}

int pen(int N, int L, const isw_ocean* o, const double* h, double* T, const double* qocean, const double* qsw,
        double dt, double* qns_acc, double* sun_acc, const double* land, const double* up, hipStream_t st) {
  const double bytes = (double)N * 8 * (L + 2 * L + 2 + 4 + 1 + (up ? 1 : 0));
  ISW_BY_L(L, ISW_LAUNCH("ocn_pen", bytes, k_ocn_pen<LL>, grid1(N), dim3(BS), N, h, T, qocean, qsw, o->rho_w * o->cw,
                         o->sw_R, o->sw_z1, o->sw_z2, dt, qns_acc, sun_acc, land, up));
  return (int)hipGetLastError();
}

int fresh(int N, const isw_ocean* o, const double* h, double* S, const double* evap, const double* rain,
          const double* ice, double* ice_prev, double dt, double* fw_acc, const double* land, double* mouth,
          double* hw, double* M, const double* area, hipStream_t st) {
  const double bytes = (double)N * 8 * (1 + 2 + 3 + 2 + 2 + 1 + (mouth ? 2 : 0) + (o->fw_local > 0 ? 5 : 0));
  ISW_LAUNCH("ocn_fresh", bytes, k_ocn_fresh, grid1(N), dim3(BS), N, h, S, evap, rain, ice, ice_prev, o->S_ref,
             o->rho_w, dt, fw_acc, land, mouth, o->fw_local, hw, M, area);
  return (int)hipGetLastError();
}

int diffuse(const isw_grid* g, int L, const double* h, const double* T, const double* S, double* Tn, double* Sn,
            const double* wall, double kappa_h, double dt, hipStream_t st) {
  const int N = g->N;
  const double bytes = (double)N * 8 * (5.0 * L) + (double)N * 6 * 4 + (double)g->E * (2 * 4 + 2 * 8) + N * 8.0;
  ISW_LAUNCH("ocn_diffuse", bytes, k_ocn_diffuse, dim3((unsigned)((N + BS - 1) / BS), (unsigned)L), dim3(BS), N, h,
             T, S, Tn, Sn, g->cellEdge, g->eI, g->eJ, g->le, g->de, g->areaCell, wall, kappa_h, dt);
  return (int)hipGetLastError();
}

int forcing(const isw_grid* g, double dt, double* tau_acc, double* taum, double* q_acc, double* sun_acc,
            double* fw_acc, hipStream_t st) {
  const int N = g->N, E = g->E;
  ISW_LAUNCH("ocn_rms", 2.0 * E * 8, k_ocn_rms, grid1(E), dim3(BS), E, 1.0 / dt, tau_acc);
  const double bytes = (double)E * 8 * 3 + (double)N * (6 * 4 + 8 + 8 + 3 * 16);
  ISW_LAUNCH("ocn_means", bytes, k_ocn_means, grid1(N), dim3(BS), N, tau_acc, g->cellEdge, g->le, g->de, g->areaCell,
             taum, dt, q_acc, sun_acc, fw_acc);
  return (int)hipGetLastError();
}

}  // namespace isw_ocn
