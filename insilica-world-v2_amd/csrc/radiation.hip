// The k-table radiation as gfx950 HIP kernels: the longwave's sub-level set-up and column pass, the shortwave's
// column pass (reference: src/insilica/radiation.py; kernels/air.cu:2111-2511).
//
// Compiled twice, as core.hip: ISW_FORM_MIXED=0 with -ffp-contract=off into isw_f64 (the plain form, `fast` = double,
// the reference's double build with its float32 storage points kept: tables, sub-levels, clamped table positions and
// per-g-point results), ISW_FORM_MIXED=1 with -ffp-contract=fast into isw_mix (`fast` = float and the transcendentals
// in float, as the reference's MIXED build; every g-point sum stays FP64).
//
// Design (MI355X): the column pass is one 64-aligned block per column, one thread per g-point, the column's sub-level
// properties staged once in LDS. Kernels are instantiated per layer count so that every per-g-point array (optical
// depths; the shortwave's reflectances, transmittances and albedos) has a compile-time size and stays in VGPRs. The
// shortwave computes its sub-levels from the bottom up, so the clear sky's adding and the all-sky albedos run in the
// same loop (the clear sky needs nothing stored); between its two passes it keeps the albedos (and in FP64 the gas
// depths), recomputing each sub-level's optics on the way down. Per-g results go to
// LDS as float32 (the reference's red[][] width), then every wave sums whole rows: each lane adds its g-points in
// order, in FP64, and a butterfly finishes the row -- a fixed order, so results repeat bit for bit. Blocks are
// renumbered so that each XCD takes a contiguous range of columns (the sub-level reads of neighbouring columns share
// cache lines).
#include <hip/hip_runtime.h>

#include <cmath>
#include <type_traits>

#include "diag.hpp"
#include "insilica_hip.h"

#if ISW_FORM_MIXED
#define NS isw_mix
typedef float fast;
#else
#define NS isw_f64
typedef double fast;
#endif

namespace NS {

constexpr int NSUB = 4, NVP = 15, NGMAX = 256;
constexpr double PI_ = 3.14159265358979323846;

// The reference's transcendental macros (air.cu:26-38): FP64 in the plain form; in the mixed form FEXP and FPOW run
// in float and return double.
__device__ __forceinline__ double FEXPd(double x) {
#if ISW_FORM_MIXED
  return (double)expf((float)x);
#else
  return exp(x);
#endif
}
__device__ __forceinline__ double FPOWd(double x, double y) {
#if ISW_FORM_MIXED
  return (double)powf((float)x, (float)y);
#else
  return pow(x, y);
#endif
}
// The column passes' exp and division: FP64 in the plain form (the reference's operations); in the mixed form the
// hardware's: exp as 2^(x log2 e) (v_exp_f32) and a / b as a times the hardware reciprocal, each within a few float
// steps (the mixed form's bar against FP64 is 1e-4 of the fluxes).
__device__ __forceinline__ fast fexp(fast x) {
#if ISW_FORM_MIXED
  return __expf(x);
#else
  return exp(x);
#endif
}
__device__ __forceinline__ fast fdiv(fast a, fast b) {
#if ISW_FORM_MIXED
  return __fdividef(a, b);
#else
  return a / b;
#endif
}

// XCD-aware block order (as the stencils'): XCD x takes one contiguous range of blocks.
__device__ __forceinline__ int xblock() {
  const int nb = (int)gridDim.x, b = (int)blockIdx.x;
  const int q = nb >> 3, r = nb & 7, x = b & 7, k = b >> 3;
  return (x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + k;
}

// The parameters the radiation reads from the reference's PAR_ORDER array (air.cu:118-127), by value.
struct RadP {
  double eps_s, grav, cp, kappa, p0, tilt, year_s, day_s, cloud_g, bl_q, ice_air, lw_refl, ecc, lam_peri;
  int edge_k;
};

static RadP make_radp(const double* a) {
  RadP p;
  p.eps_s = a[4]; p.grav = a[13]; p.cp = a[14]; p.kappa = a[15]; p.p0 = a[16]; p.tilt = a[17]; p.year_s = a[18];
  p.day_s = a[19]; p.cloud_g = a[59]; p.bl_q = a[89]; p.edge_k = (int)a[101]; p.ice_air = a[102];
  p.lw_refl = a[124]; p.ecc = a[126]; p.lam_peri = a[127];
  return p;
}

// A table position clamped as the reference writes it, fminf(fmaxf(x, 0.f), (n - 1) - 1e-4f): float arguments,
// so the position is rounded to float in either form.
__device__ __forceinline__ float clampf_pos(fast x, int n) {
  return fminf(fmaxf((float)x, 0.f), (float)((fast)(n - 1) - (fast)1e-4f));
}

// Tetens over water (air.cu:160-162), and blended with ice below freezing (air.cu:165-171).
__device__ __forceinline__ fast qsat_w(fast T, fast p, double eps) {
  return (fast)eps * (fast)611.0 * (fast)FEXPd((fast)17.27 * (T - (fast)273.15) / (T - (fast)35.85)) / p;
}
__device__ __forceinline__ fast qsat_a(fast T, fast p, double ice, double eps) {
  const fast w = qsat_w(T, p, eps);
  if (!(ice > 0) || T >= (fast)273.15) return w;
  const fast l = fmin((fast)1, fmax((fast)0, (T - (fast)253.15) / (fast)20));
  const fast si = (fast)eps * (fast)611.0 * (fast)FEXPd((fast)21.875 * (T - (fast)273.16) / (T - (fast)7.66)) / p;
  return l * w + (1 - l) * si;
}

// ---------------------------------------------------------------------------------------------------------------
// lw_setup (air.cu:2145-2229): one thread per (column, layer); 64 columns per block, blockIdx.y the layer.
// Per sub-level: edge temperatures (the layer isothermal above the 380 K edge, else T ~ p^n keeping the layer's mean
// temperature), table positions, gas amounts (vapor at the layer's mean relative humidity), cloud cover and optical
// depth (in the sub-level holding the median of the layer's vapor), air molecules, the haze's share, CIA weight.
// ---------------------------------------------------------------------------------------------------------------
template <int L>
__global__ void __launch_bounds__(64) k_lw_setup(int N, RadP P, isw_lw_tables t, isw_air air, double co2,
                                                  double n2o, double ch4, const double* __restrict__ gx,
                                                  const double* __restrict__ mu_in, const double* __restrict__ theta,
                                                  const double* __restrict__ qv, const double* __restrict__ cloudo,
                                                  double hs, int tmode, float* __restrict__ sub) {
  // No contraction here in either form: the layer edges, the sub-levels' pressures and the vapor's running share are
  // then the same bits in both, and so is every discrete choice made from them (the cloud's sub-level: an even vapor
  // profile puts the median exactly on a sub-level boundary, decided by the last bit). The kernel is a small share of
  // the radiation's time.
#pragma clang fp contract(off)
  const int i = xblock() * 64 + (int)threadIdx.x, k = (int)blockIdx.y;
  if (i >= N) return;
  double pb[L], pib[L], th_k = 0, q_k = 0;
  {
    double p = 0, xt = 0;
#pragma unroll
    for (int j = 0; j < L; j++) {
      const double pt = p;
      p = p + P.grav * mu_in[(size_t)j * N + i];
      pb[j] = p;
      const double x = pow(p / P.p0, P.kappa);
      pib[j] = P.cp * (p * x - pt * xt) / ((P.kappa + 1) * (p - pt));
      xt = x;
    }
  }
  th_k = theta[(size_t)k * N + i];
  q_k = qv ? fmax(qv[(size_t)k * N + i], 0.0) : 0.0;
  double pt = 0, pbk = 0, pibk = 0, pbL = pb[L - 1];
#pragma unroll
  for (int j = 0; j < L; j++)
    if (j == k) { pbk = pb[j]; pibk = pib[j]; }
#pragma unroll
  for (int j = 0; j < L; j++)
    if (j == k - 1) pt = pb[j];
  const bool strat = k < P.edge_k && L >= 3;
  const double Tiso = pibk * th_k / P.cp;
  const double nT = tmode == 1 ? P.kappa : 0.190;
  const double Cn = Tiso * (nT + 1) * (pbk - pt) / (pow(pbk, nT + 1) - pow(pt, nT + 1));
  auto Tof = [&](double p) -> double { return tmode == 0 || strat ? Tiso : Cn * FPOWd(p, nT); };
  const fast fcl = cloudo ? (fast)cloudo[(size_t)k * N + i] : (fast)0;
  const fast tcl = cloudo ? (fast)cloudo[(size_t)(L + 3 + k) * N + i] : (fast)0;
  const double FT[5] = {0, 0.05, 0.2, 0.5, 1}, FE[5] = {0, 0.25, 0.5, 0.75, 1};
  // the layer's vapor sum and the cloud's median test stay FP64 in both forms: with the vapor even (above the edge) the
  // median falls exactly on a sub-level boundary, where float rounding would move the cloud to the next sub-level
  fast qs[NSUB], uas[NSUB];
  double qm = 0;
  double pe[NSUB + 1];
#pragma unroll
  for (int j = 0; j <= NSUB; j++) pe[j] = pt + (k == 0 ? FT[j] : FE[j]) * (pbk - pt);
  const bool blq = P.bl_q > 0 && k == L - 1 && k > 0;
#define W(s, v) sub[((size_t)(s) * NVP + (v)) * N + i]
#pragma unroll
  for (int j = 0; j < NSUB; j++) {
    const int s = k * NSUB + j;
    const double a = pe[j], b = pe[j + 1], pm = a > 0 ? (b - a) / log(b / a) : 0.5 * b;
    const double Tm = Tof(0.5 * (a + b));
    W(s, 0) = (float)Tof(a);
    W(s, 1) = (float)Tof(b);
    // (FP64 pow in both forms: the 160 K cut-off below decides whether a sub-level holds vapor, which moves the cloud)
    const double Tq = th_k * pow(0.5 * (a + b) / P.p0, P.kappa);
    qs[j] = strat || blq ? (fast)1 : (Tq < 160 ? (fast)0 : qsat_a((fast)Tq, (fast)pm, P.ice_air, air.eps));
    qm += (double)qs[j] * ((b - a) / (pbk - pt));
    const fast x = (fast)((t.lp0 - log(pm * air.broad)) / t.dlp), y = (fast)((Tm - t.t0) / t.dtt);
    W(s, 2) = clampf_pos(x, t.NP);
    W(s, 3) = clampf_pos(y, t.NT);
    const fast ua = (fast)((b - a) / (P.grav * air.m_air_kg) * 6.02214076e23 / 1e4);
    const fast yo = clampf_pos((fast)((t.o3lp0 - log(pm)) / t.o3dlp), t.NO3);
    const int io = (int)yo;
    const fast xo3 = (fast)((t.o3[io] + (t.o3[io + 1] - t.o3[io]) * (yo - (fast)io)) * (fast)air.o3);
    const double xc = gx ? gx[(size_t)k * N + i] : co2, xn = gx ? gx[(size_t)(L + k) * N + i] : n2o;
    const double xm = gx ? gx[(size_t)(2 * L + k) * N + i] : ch4;
    W(s, 5) = (float)((fast)xc * ua);
    W(s, 6) = (float)(xo3 * ua);
    W(s, 7) = (float)((fast)xn * ua);
    W(s, 8) = (float)((fast)xm * ua);
    W(s, 11) = (float)fcl;
    W(s, 12) = (float)ua;
    W(s, 13) = (float)(FPOWd(b / pbL, 8000.0 / hs) - FPOWd(a / pbL, 8000.0 / hs));
    W(s, 14) = air.cia ? (float)((double)ua * 1e-20 * (pm / (1.380649e-23 * Tm) * 1e-6) * 1e-20) : 0.f;
    uas[j] = ua;
  }
  // the cloud's sub-level: the one holding the median of the layer's vapor
  int jc = NSUB - 1;
  {
    double cum = 0;
    bool found = false;
#pragma unroll
    for (int j = 0; j < NSUB; j++) {
      cum += qm > 0 ? (double)qs[j] * ((pe[j + 1] - pe[j]) / (pbk - pt)) / qm : 0.25;
      if (!found && cum >= 0.5) { jc = j; found = true; }
    }
  }
  if (blq) jc = 0;
#pragma unroll
  for (int j = 0; j < NSUB; j++) {
    const int s = k * NSUB + j;
    W(s, 10) = j == jc ? (float)tcl : 0.f;
    const fast qj = qm > 0 ? (fast)(q_k * (double)qs[j] / qm) : (fast)q_k;
    const fast x = qj / (1 - qj) * (fast)(air.m_air / 18.01528);
    W(s, 4) = (float)(x * uas[j]);
    W(s, 9) = (float)(x / (fast)0.04);
  }
#undef W
}

// ln k bilinear in (log p, T) on one table cell, then exp (air.cu:2266-2271).
__device__ __forceinline__ fast kbil(const float* __restrict__ K, int NT, int NG, fast fp, fast ft) {
  const fast v00 = K[0], v01 = K[NG], v10 = K[(size_t)NT * NG], v11 = K[(size_t)NT * NG + NG];
  return fexp(v00 + (v10 - v00) * fp + (v01 - v00) * ft + (v00 - v01 - v10 + v11) * fp * ft);
}

// Two-stream through one sub-level, far side's Planck radiance Bf to the near side's Bn, linear in optical depth,
// diffusivity 1.66 (air.cu:2123-2127).
__device__ __forceinline__ fast lw_through(fast I, fast tau, fast Bn, fast Bf) {
  const fast x = tau * (fast)1.66, tr = fexp(-x);
  const fast lin = x > (fast)1e-3 ? 1 - fdiv(1 - tr, x) : x * ((fast)0.5 - x / 6);
  return I * tr + Bf * (1 - tr) + (Bn - Bf) * lin;
}

// Rows of per-g float values in LDS (rows x ngp) summed over g-points: each wave takes whole rows, each lane adds its
// g-points in order in FP64, a butterfly finishes the row; sums[r] gets the result (every row, fixed order).
template <int ROWS>
__device__ __forceinline__ void g_sums(const float* red, int ngp, double* sums) {
  const int lane = (int)threadIdx.x & 63, w = (int)threadIdx.x >> 6, nw = ngp >> 6;
  for (int r = w; r < ROWS; r += nw) {
    double v = 0;
    for (int m = 0; m < nw; m++) v += (double)red[r * ngp + m * 64 + lane];
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o);
    if (lane == 0) sums[r] = v;
  }
}

// ---------------------------------------------------------------------------------------------------------------
// lw_column (air.cu:2231-2348): one block per column, one thread per g-point (block = NG rounded up to 64).
// ---------------------------------------------------------------------------------------------------------------
template <int L>
__global__ void __launch_bounds__(NGMAX) k_lw_column(int N, RadP P, isw_lw_tables t, const float* __restrict__ sub,
                                                      const double* __restrict__ Ts, double* __restrict__ radf) {
  constexpr int NSS = NSUB * L, RW = L + 1, E0 = 3 * RW + 3, ROWS = E0 + 6 * L;
  extern __shared__ float red[];                       // ROWS x ngp
  __shared__ fast sTt[NSS], sTb[NSS], sfp[NSS], sft[NSS], su[NSS][5], sfh[NSS], stc[NSS], sfc[NSS], scia[NSS];
  __shared__ int sip[NSS], sit[NSS];
  __shared__ double sums[ROWS];
  const int i = xblock(), g = (int)threadIdx.x, ngp = (int)blockDim.x;
  const int NG = t.NG, NT = t.NT, NTP = t.NTP;
  if (g < NSS) {
    const int s = g;
#define R_(v) sub[((size_t)s * NVP + (v)) * N + i]
    sTt[s] = R_(0);
    sTb[s] = R_(1);
    // lw_setup writes the positions inside the tables; clamped again here so that no input reads outside them
    const fast x = clampf_pos(R_(2), t.NP), y = clampf_pos(R_(3), NT);
    sip[s] = (int)x;
    sfp[s] = x - (fast)sip[s];
    sit[s] = (int)y;
    sft[s] = y - (fast)sit[s];
#pragma unroll
    for (int c = 0; c < 5; c++) su[s][c] = R_(4 + c);
    sfh[s] = R_(9);
    stc[s] = R_(10);
    sfc[s] = R_(11);
    scia[s] = R_(14);
#undef R_
  }
  __syncthreads();
#define RED(r, v) red[(r) * ngp + g] = (float)(v)
  if (g < NG) {
    const fast refl = P.lw_refl > 0 ? (fast)(1 - P.eps_s) : (fast)0;
    const size_t cs = (size_t)t.NP * NT * NG;
    fast tau[NSS];
#pragma unroll
    for (int s = 0; s < NSS; s++) {
      const float* K = t.logk + ((size_t)sip[s] * NT + sit[s]) * NG + g;
      const fast fp = sfp[s], ft = sft[s];
      fast lk[6];
#pragma unroll
      for (int c = 0; c < 6; c++) lk[c] = kbil(K + c * cs, NT, NG, fp, ft);
      tau[s] = fmax((fast)0, su[s][0] * (lk[0] + (lk[1] - lk[0]) * sfh[s]) + su[s][1] * lk[2] + su[s][2] * lk[3] +
                                 su[s][3] * lk[4] + su[s][4] * lk[5]);
      if (t.cia) {
        const float* C = t.cia + (size_t)sit[s] * NG + g;
        tau[s] += scia[s] * (C[0] + (C[NG] - C[0]) * ft);
      }
    }
    const float* pg = t.planck;
    auto Bof = [&](fast T) -> fast {
      const float y = fminf(fmaxf((float)(T - (fast)t.tp0), 0.f), (float)((fast)(NTP - 1) - (fast)1e-4f));
      const int j = (int)y;
      const float b0 = pg[(size_t)j * NG + g], b1 = pg[(size_t)(j + 1) * NG + g];
      return b0 + (fast)(b1 - b0) * ((fast)y - (fast)j);
    };
    const fast PIf = (fast)PI_;
    fast I = 0, Icl = 0;
    RED(RW, 0.f);
#pragma unroll
    for (int k = 0; k < L; k++) {
      const fast f = sfc[k * NSUB];
      fast Ic = I, Id = I;
#pragma unroll
      for (int s = k * NSUB; s < (k + 1) * NSUB; s++) {
        const fast Bt = Bof(sTt[s]), Bb = Bof(sTb[s]);
        Ic = lw_through(Ic, tau[s], Bb, Bt);
        if (refl > 0) Icl = lw_through(Icl, tau[s], Bb, Bt);
        if (f > 0) Id = lw_through(Id, tau[s] + stc[s] / (fast)1.66, Bb, Bt);
        if (s == k * NSUB) RED(E0 + 6 * k + 1, PIf * (f > 0 ? (1 - f) * Ic + f * Id : Ic));
        if (s == (k + 1) * NSUB - 2) RED(E0 + 6 * k + 4, PIf * (f > 0 ? (1 - f) * Ic + f * Id : Ic));
      }
      I = f > 0 ? (1 - f) * Ic + f * Id : Ic;
      RED(RW + k + 1, PIf * I);
    }
    const fast Bs = Bof((fast)Ts[i]);
    fast Iu = refl * I, tr = 1, Iuc = refl * Icl, trc = 1;
    RED(L, PIf * Iu);
    RED(2 * RW + L, Bs);
#pragma unroll
    for (int k = L - 1; k >= 0; k--) {
      const fast f = sfc[k * NSUB];
      fast Ic = Iu, Id = Iu, tc = tr, td = tr;
#pragma unroll
      for (int s = (k + 1) * NSUB - 1; s >= k * NSUB; s--) {
        const fast Bt = Bof(sTt[s]), Bb = Bof(sTb[s]);
        const fast t1 = fexp(-tau[s] * (fast)1.66);
        Iuc = lw_through(Iuc, tau[s], Bt, Bb);
        trc *= t1;
        Ic = lw_through(Ic, tau[s], Bt, Bb);
        tc *= t1;
        if (f > 0) {
          Id = lw_through(Id, tau[s] + stc[s] / (fast)1.66, Bt, Bb);
          td *= t1 * fexp(-stc[s]);
        }
        const int e = s == (k + 1) * NSUB - 1 ? 3 : (s == k * NSUB + 1 ? 0 : -1);
        if (e >= 0) {
          RED(E0 + 6 * k + e, PIf * (f > 0 ? (1 - f) * Ic + f * Id : Ic));
          RED(E0 + 6 * k + e + 2, (f > 0 ? (1 - f) * tc + f * td : tc) * Bs);
        }
      }
      Iu = f > 0 ? (1 - f) * Ic + f * Id : Ic;
      tr = f > 0 ? (1 - f) * tc + f * td : tc;
      RED(k, PIf * Iu);
      RED(2 * RW + k, tr * Bs);
    }
    RED(3 * RW, PIf * Iuc);
    RED(3 * RW + 1, trc * Bs);
    RED(3 * RW + 2, Bs);
  } else {
    for (int r = 0; r < ROWS; r++) RED(r, 0.f);
  }
#undef RED
  __syncthreads();
  g_sums<ROWS>(red, ngp, sums);
  __syncthreads();
  if (g < ROWS) {
    double v = sums[g];
    if ((g >= 2 * RW && g < 3 * RW) || g == 3 * RW + 1 || (g >= E0 && (g - E0) % 3 == 2)) v /= sums[3 * RW + 2];
    radf[(size_t)g * N + i] = v;
  }
}

// ---------------------------------------------------------------------------------------------------------------
// sw_layer (air.cu:2358-2382): Delta-Eddington reflectance and transmittance of one homogeneous sub-level, for diffuse
// light (Rd, Td) and for the sun's beam at cosine mu0 (Rs, Ts; Tdd the beam's direct transmission).
// ---------------------------------------------------------------------------------------------------------------
__device__ __forceinline__ void sw_layer(fast tau, fast ssa, fast g, fast mu0, fast& Rd, fast& Td, fast& Rs, fast& Ts,
                                         fast& Tdd) {
  const fast f = g * g;
  tau = (1 - ssa * f) * tau;
  ssa = fdiv((1 - f) * ssa, 1 - ssa * f);
  g = fdiv(g - f, 1 - f);
  const fast g1 = (7 - ssa * (4 + 3 * g)) / 4, g2 = -(1 - ssa * (4 - 3 * g)) / 4, g3 = (2 - 3 * g * mu0) / 4;
  const fast g4 = 1 - g3;
  const fast a1 = g1 * g4 + g2 * g3, a2 = g1 * g3 + g2 * g4;
  const fast k = sqrt(fmax((g1 - g2) * (g1 + g2), (fast)1e-12));
  const fast e = fexp(-k * tau), e2 = e * e, k2e = 2 * k * e;
  fast rf = fdiv(1, k + g1 + (k - g1) * e2);
  Rd = g2 * (1 - e2) * rf;
  Td = k2e * rf;
  Tdd = fexp(fdiv(-tau, mu0));
  const fast km = k * mu0;
  fast om = 1 - km * km;
  if (fabs(om) < (fast)1e-3) om = om < 0 ? (fast)-1e-3 : (fast)1e-3;
  rf = fdiv(ssa * rf, om);
  Rs = rf * ((1 - km) * (a2 + k * g3) - (1 + km) * (a2 - k * g3) * e2 - k2e * (g3 - a2 * mu0) * Tdd);
  Ts = rf * (k2e * (g4 + a1 * mu0) - Tdd * ((1 + km) * (a1 + k * g4) - (1 - km) * (a1 - k * g4) * e2));
  Rs = fmin(fmax(Rs, (fast)0), 1 - Tdd);
  Ts = fmin(fmax(Ts, (fast)0), 1 - Tdd - Rs);
}

// The sun's longitude at time t (air.cu:174-185): Kepler's equation by five Newton steps (circular orbit: uniform).
__device__ double sun_lon(double t, const RadP& P) {
  if (!(P.ecc > 0)) return 2 * PI_ * t / P.year_s;
  const double e = P.ecc, nu0 = -P.lam_peri;
  const double E0 = 2 * atan(sqrt((1 - e) / (1 + e)) * tan(nu0 / 2));
  const double M = E0 - e * sin(E0) + 2 * PI_ * t / P.year_s;
  double E = M;
  for (int it = 0; it < 5; it++) E -= (E - e * sin(E) - M) / (1 - e * cos(E));
  return 2 * atan2(sqrt(1 + e) * sin(E / 2), sqrt(1 - e) * cos(E / 2)) + P.lam_peri;
}

// ---------------------------------------------------------------------------------------------------------------
// sw_column (air.cu:2384-2511): one block per column, one thread per g-point.
// ---------------------------------------------------------------------------------------------------------------
template <int L>
__global__ void __launch_bounds__(NGMAX) k_sw_column(int N, RadP P, isw_sw_tables t, isw_air air,
                                                      const float* __restrict__ sub, int mode,
                                                      const double* __restrict__ tnow, double interval,
                                                      const double* __restrict__ pos, const double* __restrict__ alb,
                                                      const double* __restrict__ cloudo, double* __restrict__ swf) {
  constexpr int NSS = NSUB * L, ROWS = 3 * L + 2;
  extern __shared__ float red[];                       // ROWS x ngp
  __shared__ fast su[NSS][6], sfh[NSS], sfp[NSS], sft[NSS], sr[NSS], sh[NSS], stc[NSS], sfc[NSS], sli[NSS], scia[NSS];
  __shared__ int sip[NSS], sit[NSS];
  __shared__ fast smu;
  __shared__ double sums[ROWS];
  const int i = xblock(), g = (int)threadIdx.x, ngp = (int)blockDim.x;
  const int NG = t.NG, NT = t.NT;
  if (g < NSS) {
    const int s = g, k = s / NSUB;
#define R_(v) sub[((size_t)s * NVP + (v)) * N + i]
    // (positions clamped to this table's axes: the same values for lw_setup's sub-levels when the tables share their
    // pressure axis, and no read outside the table for any input)
    const fast x = clampf_pos(R_(2), t.NP);
    const fast y = fmaxf(fminf((fast)R_(3) * (fast)t.yscale, (fast)(NT - 1) - 1e-4f), 0.f);
    sip[s] = (int)x;
    sfp[s] = x - (fast)sip[s];
    sit[s] = (int)y;
    sft[s] = y - (fast)sit[s];
#pragma unroll
    for (int c = 0; c < 5; c++) su[s][c] = R_(4 + c);
    su[s][5] = (fast)air.o2 * R_(12);
    sfh[s] = R_(9);
    sr[s] = R_(12);
    sh[s] = R_(13);
    scia[s] = R_(14);
    const bool cl = R_(10) > 0 && cloudo;
    sfc[s] = cl ? (fast)cloudo[(size_t)k * N + i] : (fast)0;
    stc[s] = cl ? (fast)(t.inhom * cloudo[(size_t)(2 * L + 3 + k) * N + i]) : (fast)0;
    sli[s] = cl && t.cssa_i ? (fast)cloudo[(size_t)(3 * L + 3 + k) * N + i] : (fast)1;
#undef R_
  }
  if (g == ngp - 1) {                                 // the sun: the mean of its cosine over the interval's daylight
    fast m = 0.5;
    if (mode == 0) {
      fast sum = 0;
      int lit = 0;
      for (int j = 0; j < 4; j++) {
        const double tt = tnow[0] + (j + 0.5) * interval / 4;
        const double ang = sun_lon(tt, P), dec = asin(sin(P.tilt) * sin(ang));
        const double rot = 2 * PI_ * fmod(tt, P.day_s) / P.day_s;
        const double cz = pos[3 * i] * cos(dec) * cos(-rot) + pos[3 * i + 1] * cos(dec) * sin(-rot) +
                          pos[3 * i + 2] * sin(dec);
        if (cz > 0) {
          sum += (fast)cz;
          lit++;
        }
      }
      m = lit ? sum / lit : (fast)0.05;
    }
    smu = fmax(m, (fast)0.02);
  }
  __syncthreads();
#define RED(r, v) red[(r) * ngp + g] = (float)(v)
  if (g < NG) {
    const fast mu0 = smu, a = (fast)alb[i];
    const size_t cs = (size_t)t.NP * NT * NG;
    const fast cssa = t.cssa[g], ray = t.ray[g], hzf = t.hzf[g];
    const fast cssa_i = t.cssa_i ? (fast)t.cssa_i[g] : (fast)0, g_i = t.g_i ? (fast)t.g_i[g] : (fast)0;
    // Sub-level s's optics from its gas optical depth tg: the clear sky's (c*) and the all-sky's (the cloud's share
    // mixed in), as the reference computes them (air.cu:2440-2458). Evaluated twice per sub-level, in the albedos'
    // pass up and the fluxes' pass down, so that only tg and the albedos are kept between the passes (a third of the
    // registers that keeping every reflectance and transmittance would need; the table lookups are not repeated).
    auto optics = [&](int s, fast tg, fast& cRd, fast& cTd, fast& cRs, fast& cTs, fast& cTdd, fast& rd, fast& td,
                      fast& rs, fast& ts, fast& tdd) {
      const fast trs = ray * sr[s] * (fast)air.ray, ths = (fast)t.tau550 * hzf * sh[s];
      const fast tt = tg + trs + ths, sc = trs + (fast)t.hssa * ths;
      const fast ssa = tt > 0 ? fmin(fdiv(sc, tt), (fast)(1 - 1e-6)) : (fast)0;
      const fast ga = sc > 0 ? fdiv((fast)t.hg * (fast)t.hssa * ths, sc) : (fast)0;
      sw_layer(tt, ssa, ga, mu0, cRd, cTd, cRs, cTs, cTdd);
      rd = cRd; td = cTd; rs = cRs; ts = cTs; tdd = cTdd;
      const fast f = sfc[s];
      if (f > 0 && stc[s] > 0) {
        const fast tc = stc[s], t2 = tt + tc;
        fast cw = cssa, cg = (fast)P.cloud_g;
        if (t.cssa_i) {
          const fast sl = sli[s], wl = sl * cssa, wi = (1 - sl) * cssa_i;
          cw = wl + wi;
          cg = cw > 0 ? fdiv(wl * (fast)P.cloud_g + wi * g_i, cw) : cg;
        }
        const fast sc2 = sc + cw * tc;
        const fast g2 = fdiv(ga * sc + cg * cw * tc, sc2);
        fast r1, t1, r2, t3, d1;
        sw_layer(t2, fmin(fdiv(sc2, t2), (fast)(1 - 1e-6)), g2, mu0, r1, t1, r2, t3, d1);
        rd = (1 - f) * cRd + f * r1;
        td = (1 - f) * cTd + f * t1;
        rs = (1 - f) * cRs + f * r2;
        ts = (1 - f) * cTs + f * t3;
        tdd = (1 - f) * cTdd + f * d1;
      }
    };
    // sub-level s's gas optical depth from the tables (air.cu:2425-2437)
    auto gas = [&](int s) -> fast {
      const float* K = t.logk + ((size_t)sip[s] * NT + sit[s]) * NG + g;
      const fast fp = sfp[s], ft = sft[s];
      fast lk[7];
#pragma unroll
      for (int c = 0; c < 7; c++) lk[c] = kbil(K + c * cs, NT, NG, fp, ft);
      fast tg = fmax((fast)0, su[s][0] * (lk[0] + (lk[1] - lk[0]) * sfh[s]) + su[s][1] * lk[2] + su[s][2] * lk[3] +
                                  su[s][3] * lk[4] + su[s][4] * lk[5] + su[s][5] * lk[6]);
      if (t.cia) {
        const float* C = t.cia + (size_t)sit[s] * NG + g;
        tg += scia[s] * (C[0] + (C[NG] - C[0]) * ft);
      }
      return tg;
    };
    // bottom up: the gas optical depths, the clear sky's albedos (nothing kept) and the all-sky's. The FP64 form keeps
    // each sub-level's gas depth for the way down; the mixed form looks the tables up again instead (measured at
    // 7 layers: 243 VGPRs and two waves per SIMD against one, the pass 6.9 ms against 10.2 ms; in FP64 keeping wins)
#if ISW_FORM_MIXED
#define ISW_GAS_DOWN(s) gas(s)
    fast Ad[NSS + 1], As[NSS + 1];
#else
#define ISW_GAS_DOWN(s) tgs[s]
    fast tgs[NSS], Ad[NSS + 1], As[NSS + 1];
#endif
    Ad[NSS] = a;
    As[NSS] = a;
    fast cAd = a, cAs = a;
#pragma unroll
    for (int s = NSS - 1; s >= 0; s--) {
      const fast tg = gas(s);
#if !ISW_FORM_MIXED
      tgs[s] = tg;
#endif
      fast cRd, cTd, cRs, cTs, cTdd, rd, td, rs, ts, tdd;
      optics(s, tg, cRd, cTd, cRs, cTs, cTdd, rd, td, rs, ts, tdd);
      {
        const fast den = fdiv(1, 1 - cAd * cRd);
        const fast nAd = cRd + cTd * cTd * cAd * den;
        cAs = cRs + (cTdd * cAs + cTs * cAd) * cTd * den;
        cAd = nAd;
      }
      const fast den = fdiv(1, 1 - Ad[s + 1] * rd);
      Ad[s] = rd + td * td * Ad[s + 1] * den;
      As[s] = rs + (tdd * As[s + 1] + ts * Ad[s + 1]) * td * den;
    }
    // top down: the fluxes, each layer's absorption, its top and bottom sub-level's, the surface's
    const fast wg = t.sun[g];
    fast Fdir = 1, Fdif = 0, net = 1 - As[0], absk = 0;
#pragma unroll
    for (int s = 0; s < NSS; s++) {
      fast cRd, cTd, cRs, cTs, cTdd, rd, td, rs, ts, tdd;
      optics(s, ISW_GAS_DOWN(s), cRd, cTd, cRs, cTs, cTdd, rd, td, rs, ts, tdd);
#undef ISW_GAS_DOWN
      const fast den = fdiv(1, 1 - Ad[s + 1] * rd);
      const fast Fdn = Fdir * tdd;
      Fdif = (td * Fdif + rd * As[s + 1] * Fdn + ts * Fdir) * den;
      Fdir = Fdn;
      const fast n2 = Fdir + Fdif - (Ad[s + 1] * Fdif + As[s + 1] * Fdir);
      const int k = s / NSUB;
      if (s % NSUB == 0) absk = 0;
      absk += net - n2;
      if (s % NSUB == 0) RED(L + 2 + 2 * k, wg * (net - n2));
      if (s % NSUB == NSUB - 1) {
        RED(L + 3 + 2 * k, wg * (net - n2));
        RED(k, wg * absk);
      }
      net = n2;
    }
    RED(L, wg * net);
    RED(L + 1, wg * cAs);
  } else {
    for (int r = 0; r < ROWS; r++) RED(r, 0.f);
  }
#undef RED
  __syncthreads();
  g_sums<ROWS>(red, ngp, sums);
  __syncthreads();
  if (g < ROWS) swf[(size_t)g * N + i] = sums[g];
}

// ---------------------------------------------------------------------------------------------------------------
// Host side: one stream, the layer count picks the instantiation.
// ---------------------------------------------------------------------------------------------------------------
static inline int ngpad(int NG) { return (NG + 63) / 64 * 64; }

#define ISW_BY_L(CALL)                                                                                   \
  switch (L) {                                                                                           \
    case 1: CALL(1); break;                                                                              \
    case 2: CALL(2); break;                                                                              \
    case 3: CALL(3); break;                                                                              \
    case 4: CALL(4); break;                                                                              \
    case 5: CALL(5); break;                                                                              \
    case 6: CALL(6); break;                                                                              \
    case 7: CALL(7); break;                                                                              \
    case 8: CALL(8); break;                                                                              \
    default: return ISW_EBADARG;                                                                         \
  }

int lw_setup(int N, int L, const double* par, const isw_lw_tables* t, const isw_air* air, double co2, double n2o,
             double ch4, const double* gx, const double* mu, const double* theta, const double* qv,
             const double* cloudo, double hs, int tmode, float* sub, hipStream_t st) {
  const RadP P = make_radp(par);
  const dim3 grid((unsigned)((N + 63) / 64), (unsigned)L), block(64);
  const double bytes = (double)L * N * 8 * (2 + (qv ? 1 : 0) + (gx ? 3 : 0)) + (cloudo ? 2.0 * L * N * 8 : 0) +
                       (double)NSUB * L * NVP * N * 4;
#define CALL(LL)                                                                                              \
  ISW_LAUNCH("lw_setup", bytes, k_lw_setup<LL>, grid, block, N, P, *t, *air, co2, n2o, ch4, gx, mu, theta, qv, \
             cloudo, hs, tmode, sub)
  ISW_BY_L(CALL)
#undef CALL
  return (int)hipGetLastError();
}

int lw_column(int N, int L, const double* par, const isw_lw_tables* t, const float* sub, const double* Ts,
              double* radf, hipStream_t st) {
  const RadP P = make_radp(par);
  const int ngp = ngpad(t->NG), rows = 3 * (L + 1) + 3 + 6 * L;
  const size_t shm = (size_t)rows * ngp * sizeof(float);
  const double bytes = (double)NSUB * L * NVP * N * 4 + (double)N * 8 + (double)rows * N * 8 +
                       (double)t->NC * t->NP * t->NT * t->NG * 4 + (double)t->NTP * t->NG * 4;
#define CALL(LL)                                                                                              \
  if (shm > 65536) hipFuncSetAttribute((const void*)k_lw_column<LL>, hipFuncAttributeMaxDynamicSharedMemorySize,    \
                                       (int)shm);                                                                   \
  ISW_LAUNCH_SHM("lw_column", bytes, k_lw_column<LL>, dim3((unsigned)N), dim3((unsigned)ngp), shm, N, P, *t, sub, Ts, radf)
  ISW_BY_L(CALL)
#undef CALL
  return (int)hipGetLastError();
}

int sw_column(int N, int L, const double* par, const isw_sw_tables* t, const isw_air* air, const float* sub, int mode,
              const double* tnow, double interval, const double* pos, const double* alb, const double* cloudo,
              double* swf, hipStream_t st) {
  const RadP P = make_radp(par);
  const int ngp = ngpad(t->NG), rows = 3 * L + 2;
  const size_t shm = (size_t)rows * ngp * sizeof(float);
  const double bytes = (double)NSUB * L * NVP * N * 4 + (double)N * 8 * (1 + (mode == 0 ? 3 : 0)) +
                       (cloudo ? 3.0 * L * N * 8 : 0) + (double)rows * N * 8 + (double)t->NC * t->NP * t->NT * t->NG * 4;
#define CALL(LL)                                                                                              \
  ISW_LAUNCH_SHM("sw_column", bytes, k_sw_column<LL>, dim3((unsigned)N), dim3((unsigned)ngp), shm, N, P, *t, *air, \
                 sub, mode, tnow, interval, pos, alb, cloudo, swf)
  ISW_BY_L(CALL)
#undef CALL
  return (int)hipGetLastError();
}

}  // namespace NS
