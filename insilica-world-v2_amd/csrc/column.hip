// The air column's Held-Suarez path and the surface-exchange kernels as gfx950 HIP kernels (reference:
// src/insilica/kernels/air.cu: hs_column 2065-2109 with the crossing of the isentropic interfaces 744-850,
// surface_layer 2015-2061, cell_speed 1978-1989, cell_speed_perot 1994-2006, gas_src 2133-2137, advance_time 1975).
//
// FP64 in both forms (compiled once, -ffp-contract=off): these are N-sized passes that cost microseconds per tick,
// and the crossing's mass moves re-divide the column with an energy residual that must stay at rounding.
//
// Design (MI355X): one thread per column, 64-thread blocks (641 blocks at level 6, over all 256 CUs), blocks in XCD
// order; the column's layers (L <= 8, a compile-time instantiation) live in registers; layer-major loads are coalesced
// across the wave.
#include <hip/hip_runtime.h>

#include <cmath>
#include <type_traits>

#include "diag.hpp"
#include "insilica_hip.h"

namespace isw_col {

constexpr int BS = 64;

__device__ __forceinline__ int xblock() {
  const int nb = (int)gridDim.x, b = (int)blockIdx.x;
  const int q = nb >> 3, r = nb & 7, x = b & 7, k = b >> 3;
  return (x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + k;
}
#define XTID ((int)(xblock() * BS + threadIdx.x))
static inline dim3 grid1(long n) { return dim3((unsigned)((n + BS - 1) / BS)); }

// The parameters the column kernels read, from the reference's PAR_ORDER array (air.cu:118-127), by value.
struct CP {
  double grav, cp, kappa, p0, floor_, theta_star, Lv, moist, mu_bl, bl_mix, bl_cap, rho_ice, ice_cover, bulk_gust,
      q_sea, z0_land;
  double bl_room;                                      // (the mixed layer's room up to its condensation level)
  double ths[ISW_LMAX];
  // the column's gas tracers (3 x L mole fractions, xfer) while it is stepped, else null: every move of its air
  // carries them (air.cu:503-507)
  double* xf;
};

static CP make_cp(const double* a) {
  CP p;
  p.grav = a[13]; p.cp = a[14]; p.kappa = a[15]; p.p0 = a[16]; p.floor_ = a[11]; p.theta_star = a[10]; p.Lv = a[23];
  p.moist = a[22]; p.mu_bl = a[9]; p.bl_mix = a[70]; p.bl_cap = a[73]; p.rho_ice = a[29]; p.ice_cover = a[30];
  p.bulk_gust = a[75]; p.q_sea = a[97]; p.z0_land = a[91];
  p.bl_room = 1e30;
  for (int j = 0; j < ISW_LMAX; j++) p.ths[j] = a[132 + j];
  p.xf = nullptr;
  return p;
}

template <int L>
__device__ __forceinline__ void layer_exner(const double* mu, const CP& P, double* pb, double* pib) {
  double p = 0, xt = 0;
#pragma unroll
  for (int k = 0; k < L; k++) {
    const double pt = p;
    p = p + P.grav * mu[k];
    pb[k] = p;
    const double x = pow(p / P.p0, P.kappa);
    pib[k] = P.cp * (p * x - pt * xt) / ((P.kappa + 1) * (p - pt));
    xt = x;
  }
}

// theta* of interface j: the lowest is theta_star, the ones above from the list (from the top).
template <int L>
__device__ __forceinline__ double tstar(int j, const CP& P) { return j == L - 2 ? P.theta_star : P.ths[j]; }

// The shape s of a layer's truncated-exponential theta density between edges a < b with mean m (air.cu:245-270):
// r(s) = 1 - 1/s + 1/(e^s - 1) = (m - a) / (b - a) by Newton's method kept inside a bracket.
__device__ double edge_shape(double a, double b, double m) {
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

// The layer's mass per kelvin at its lower (fa) and upper (fb) edge (air.cu:525-539).
__device__ void edge_density(double a, double b, double m, double& fa, double& fb) {
  if (!(b > a)) {
    fa = 0;
    fb = 0;
    return;
  }
  const double W = b - a, s = edge_shape(a, b, m);
  if (fabs(s) < 1e-6) {
    fa = (1 - s / 2) / W;
    fb = (1 + s / 2) / W;
    return;
  }
  const double e1 = expm1(s);
  fa = s / e1 / W;
  fb = s * (e1 + 1) / e1 / W;
}

// The air column's own parameters (PAR_ORDER positions) and the air's constants, by value beside CP.
struct AP {
  double S0, asw, tau0, mu_ref, eps_s, h_sens, u0, alb, c_slab, sigma, tilt, year_s, day_s, tau_lin, Lf, Tf, sw_tau,
      bl_bot, theta_rise, bulk_c, bulk_z, bulk_w, edge_heat, sfc, edge_k, firn, rho_firn, c_ice, ecc, lam_peri, rd,
      k_evap, tau_conv, alb_ice, k_ice, Rv, sat_layer, alb_ice_melt, alb_ramp, lw_w0, ice_air, dry_edge, eps;
  // the clouds'
  double rh_crit, re_liq, re_ice, cloud_g, k_lw_liq, k_lw_ice, c_auto, q_auto, v_ice, tau_erode, k_revap, mu_cloud,
      l_ice, conv_revap, downdraft, revap, revap_rhs, revap_area, liq_cold, cw_qs;
  // the land's
  double alb_dry, alb_wet, soil_beta, k_inf, pond_cover, alb_snow, alb_snow_melt, alb_glacier, snow_cover, land_exch,
      c_dry, c_wet, land_rs, c_deep, rs_max, rs_sun, revap_rhl, bare_fx, k_firn;
  // the mixed boundary layer's
  double mu_night, bl_ent, bl_reach, bl_lcl, ent_lcl, vf;
  // the spread layers' capacity, the base condensation, edge vapor, freeze-drying, the surface air's vapor, the
  // plumes' root
  double sat_prof, sat_base, edge_vap, freeze_dry, q_shape, plume_root, root_q, bl_q;
  // the deep plume's
  double plume_deep, plume_ent, ent_cloud, ent_det, l_det, storm_area, graupel_v, ice_small, ent_ifs, ent_loc, det_all,
      buoy_v, shallow;
  double ml_clear, cu_cloud;                           // (the mixed layer's cloud with bl_q)
  double cu_dth, cu_dp, cu_sub, cl_float, sub_mix;     // (the cumulus layer's)
};

static AP make_ap(const double* a, const isw_air* air) {
  AP p;
  p.S0 = a[0]; p.asw = a[1]; p.tau0 = a[2]; p.mu_ref = a[3]; p.eps_s = a[4]; p.h_sens = a[5]; p.u0 = a[6];
  p.alb = a[7]; p.c_slab = a[8]; p.sigma = a[12]; p.tilt = a[17]; p.year_s = a[18]; p.day_s = a[19];
  p.tau_lin = a[20]; p.Lf = a[24]; p.Tf = a[25]; p.sw_tau = a[34]; p.bl_bot = a[54]; p.theta_rise = a[69];
  p.bulk_c = a[74]; p.bulk_z = a[76]; p.bulk_w = a[77]; p.edge_heat = a[84]; p.sfc = a[90]; p.edge_k = a[101];
  p.firn = a[120]; p.rho_firn = a[122]; p.c_ice = a[123]; p.ecc = a[126]; p.lam_peri = a[127]; p.rd = air->rd;
  p.k_evap = a[26]; p.tau_conv = a[27]; p.alb_ice = a[28]; p.k_ice = a[31]; p.Rv = a[32]; p.sat_layer = a[33];
  p.alb_ice_melt = a[41]; p.alb_ramp = a[45]; p.lw_w0 = a[48]; p.ice_air = a[102]; p.dry_edge = a[99];
  p.eps = air->eps;
  p.rh_crit = a[56]; p.re_liq = a[57]; p.re_ice = a[58]; p.cloud_g = a[59]; p.k_lw_liq = a[60]; p.k_lw_ice = a[61];
  p.c_auto = a[62]; p.q_auto = a[63]; p.v_ice = a[64]; p.tau_erode = a[66]; p.k_revap = a[67]; p.mu_cloud = a[68];
  p.l_ice = a[79]; p.conv_revap = a[87]; p.downdraft = a[92]; p.revap = a[103]; p.revap_rhs = a[104];
  p.revap_area = a[106]; p.liq_cold = a[128]; p.cw_qs = a[129];
  p.alb_dry = a[35]; p.alb_wet = a[36]; p.soil_beta = a[37]; p.k_inf = a[38]; p.pond_cover = a[39]; p.alb_snow = a[40];
  p.alb_snow_melt = a[42]; p.alb_glacier = a[43]; p.snow_cover = a[44]; p.land_exch = a[46]; p.c_dry = a[49];
  p.c_wet = a[50]; p.land_rs = a[52]; p.c_deep = a[94]; p.rs_max = a[95]; p.rs_sun = a[96]; p.revap_rhl = a[105];
  p.bare_fx = a[115]; p.k_firn = a[121];
  p.mu_night = a[71]; p.bl_ent = a[72]; p.bl_reach = a[98]; p.bl_lcl = a[100]; p.ent_lcl = a[130]; p.vf = air->vf;
  p.sat_prof = a[51]; p.sat_base = a[53]; p.edge_vap = a[86]; p.freeze_dry = a[85]; p.q_shape = a[83];
  p.plume_root = a[88]; p.root_q = a[107]; p.bl_q = a[89];
  p.plume_deep = a[47]; p.plume_ent = a[78]; p.ent_cloud = a[93]; p.ent_det = a[108]; p.l_det = a[65];
  p.ent_ifs = a[109]; p.ent_loc = a[125]; p.det_all = a[116]; p.buoy_v = a[113]; p.shallow = a[110];
  p.storm_area = a[80]; p.graupel_v = a[81]; p.ice_small = a[82];
  p.ml_clear = a[118]; p.cu_cloud = a[119];
  p.cu_dth = a[111]; p.cu_dp = a[112]; p.cu_sub = a[114]; p.cl_float = a[117]; p.sub_mix = a[131];
  return p;
}

template <bool MIX>
__device__ __forceinline__ double omexp(double x) {
  if constexpr (MIX) return (double)-expm1f(-(float)x);
  else return 1 - exp(-(x));
}
template <bool MIX>
__device__ __forceinline__ double fpow_(double x, double y) {
  if constexpr (MIX) return (double)powf((float)x, (float)y);
  else return pow(x, y);
}

// The gray optical depth down to pressure p of a column with surface pressure ps (air.cu:853-856).
__device__ __forceinline__ double tau_p(double p, double ps, const AP& A) {
  const double s = p / ps, s2 = s * s;
  return A.tau0 * (A.tau_lin * s + (1 - A.tau_lin) * s2 * s2);
}

// The sun's ecliptic longitude at t and its distance factor (air.cu:174-185): Kepler's equation, five Newton steps.
__device__ double sun_lon_fd(double t, const AP& A, double* fd) {
  constexpr double PI_ = 3.14159265358979323846;
  if (!(A.ecc > 0)) {
    *fd = 1;
    return 2 * PI_ * t / A.year_s;
  }
  const double e = A.ecc, nu0 = -A.lam_peri;
  const double E0 = 2 * atan(sqrt((1 - e) / (1 + e)) * tan(nu0 / 2));
  const double M = E0 - e * sin(E0) + 2 * PI_ * t / A.year_s;
  double E = M;
  for (int it = 0; it < 5; it++) E -= (E - e * sin(E) - M) / (1 - e * cos(E));
  const double nu = 2 * atan2(sqrt(1 + e) * sin(E / 2), sqrt(1 - e) * cos(E / 2)), r = 1 - e * cos(E);
  *fd = 1 / (r * r);
  return nu + A.lam_peri;
}

// Saturation specific humidity (air.cu:156-171): Tetens over water; with ice_air, below freezing the blend with
// Tetens over ice by the liquid share. F is the build's `fast` type.
template <typename F>
__device__ __forceinline__ F qsat_a(F T, F p, F eps, double ice) {
  const F w = eps * (F)611.0 * (F)exp(((F)17.27 * (T - (F)273.15) / (T - (F)35.85))) / p;
  if (!(ice > 0) || T >= (F)273.15) return w;
  const F l = fmin((F)1, fmax((F)0, (T - (F)253.15) / (F)20));
  const F si = eps * (F)611.0 * (F)exp(((F)21.875 * (T - (F)273.16) / (T - (F)7.66))) / p;
  return l * w + (1 - l) * si;
}

// Saturation adjustment at Exner pib and pressure p in one Newton step (air.cu:190-198); returns the condensate.
template <bool MIX>
__device__ double condense(double& th, double& q, double pib, double p, const CP& P, const AP& A) {
  using F = std::conditional_t<MIX, float, double>;
  const F T = (F)pib * (F)th / (F)P.cp, qs = qsat_a<F>(T, (F)p, (F)A.eps, A.ice_air);
  if (!(q > qs)) return 0;
  const double dq = (F)(q - qs) / ((F)1 + (F)P.Lv * (F)P.Lv * qs / ((F)P.cp * (F)A.Rv * T * T));
  q -= dq;
  th += P.Lv * dq / pib;
  return dq;
}

// The mass share m's theta in a layer of shape s as a fraction of the way between its edges, ln(1 + m (e^s - 1)) / s
// (air.cu:233-244; es = e^s - 1). The mixed build in float32 with asymptotes beyond |s| = 30.
template <bool MIX>
__device__ __forceinline__ double spread_of(double m, double s, double es) {
  if constexpr (MIX) {
    const float mf = (float)m, sf = (float)s;
    if (fabsf(sf) < 1e-6f) return mf;
    if (sf > 30.f) return 1.f + logf(mf) / sf;
    if (sf < -30.f) return log1pf(-mf) / sf;
    return log1pf(mf * expm1f(sf)) / sf;
  } else {
    return fabs(s) < 1e-6 ? m : log1p(m * es) / s;
  }
}

// The layer's vapor capacity qs over its pressure span [pt, pb] and its change with theta dqs (air.cu:334-359):
// 4-point Gauss-Legendre, nodes colder than 160 K holding nothing, Clausius-Clapeyron's slope at each node. Its theta
// uniform, or (spread) spread between its edges ea < eb by the truncated exponential of its mean (the mass share m
// from its base at ea + (eb - ea) spread_of(m)). In the build's `fast` type, returned as such.
template <bool MIX>
__device__ __forceinline__ void layer_cap(double th, double pt, double pb, const CP& P, const AP& A,
                                          std::conditional_t<MIX, float, double>& qs,
                                          std::conditional_t<MIX, float, double>& dqs, bool spread = false,
                                          double ea = 0, double eb = 0) {
  using F = std::conditional_t<MIX, float, double>;
  constexpr double x[4] = {-0.8611363115940526, -0.3399810435848563, 0.3399810435848563, 0.8611363115940526};
  constexpr double wq[4] = {0.3478548451374538, 0.6521451548625461, 0.6521451548625461, 0.3478548451374538};
  const double s = spread ? edge_shape(ea, eb, th) : 0.0, es = spread && !MIX ? expm1(s) : 0.0;
  qs = 0;
  dqs = 0;
#pragma unroll
  for (int j = 0; j < 4; j++) {
    const double p = 0.5 * (pt + pb) + 0.5 * (pb - pt) * x[j], m = (pb - p) / (pb - pt);
    const double tn = !spread ? th : ea + (eb - ea) * spread_of<MIX>(m, s, es);
    const F ex = (F)fpow_<MIX>(p / P.p0, P.kappa), T = (F)tn * ex;
    if (T < 160) continue;
    const F v = qsat_a<F>(T, (F)p, (F)A.eps, A.ice_air);
    qs += (F)0.5 * (F)wq[j] * v;
    dqs += (F)0.5 * (F)wq[j] * v * (F)P.Lv / ((F)A.Rv * T * T) * ex;
  }
}

// Whether layer k's capacity has its theta spread between its edges (sat_prof: a middle layer below edge_k's
// tropopause, colder than its top edge), and those edges.
template <int L>
__device__ __forceinline__ bool spread_layer(int k, double th, const CP& P, const AP& A, double& ea, double& eb) {
  const bool sp = A.sat_prof > 0 && k >= (int)A.edge_k && k < L - 1 && th < tstar<L>(k - 1, P);
  ea = sp ? tstar<L>(k, P) : 0.0;
  eb = sp ? tstar<L>(k - 1, P) : 0.0;
  return sp;
}

// Layer k's capacity (the spread one with sat_prof) over [pt, pb].
template <int L, bool MIX>
__device__ __forceinline__ void layer_cap_k(int k, double th, double pt, double pb, const CP& P, const AP& A,
                                            std::conditional_t<MIX, float, double>& qs,
                                            std::conditional_t<MIX, float, double>& dqs) {
  double ea, eb;
  const bool sp = spread_layer<L>(k, th, P, A, ea, eb);
  layer_cap<MIX>(th, pt, pb, P, A, qs, dqs, sp, ea, eb);
}

// Saturation adjustment against the layer's capacity over [pt, pb] (air.cu:275-291; spread: condense_spread,
// air.cu:292-315, the whole spread shifting with the mean): one Newton step.
template <bool MIX>
__device__ double condense_layer(double& th, double& q, double pib, double pt, double pb, const CP& P, const AP& A,
                                 bool spread = false, double ea = 0, double eb = 0) {
  using F = std::conditional_t<MIX, float, double>;
  F qs, dqs;
  layer_cap<MIX>(th, pt, pb, P, A, qs, dqs, spread, ea, eb);
  if (!(q > qs)) return 0;
  const double dq = (F)(q - qs) / ((F)1 + (F)P.Lv / (F)pib * dqs);
  q -= dq;
  th += P.Lv * dq / pib;
  return dq;
}

// Condensation in layer k (air.cu:319-331): the layer's capacity (sat_layer; sat_prof: a middle layer below
// edge_k's tropopause with its theta spread between its edges; sat_base: the third layer at its base), else v1's rule
// at the pressure whose Exner function is the layer's. tk, qk: layer k's theta and vapor, updated.
template <int L, bool MIX>
__device__ double condense_at(int k, double& tk, double& qk, double pk, double pt, double mk, const CP& P,
                              const AP& A) {
  if (A.sat_layer > 0) {
    double ea, eb;
    if (spread_layer<L>(k, tk, P, A, ea, eb))
      return condense_layer<MIX>(tk, qk, pk, pt, pt + P.grav * mk, P, A, true, ea, eb);
    if (A.sat_base > 0 && k == 2) {
      const double pb = pt + P.grav * mk;
      return condense<MIX>(tk, qk, P.cp * fpow_<MIX>(pb / P.p0, P.kappa), pb, P, A);
    }
    return condense_layer<MIX>(tk, qk, pk, pt, pt + P.grav * mk, P, A);
  }
  return condense<MIX>(tk, qk, pk, fpow_<MIX>(pk / P.cp, 1 / P.kappa) * P.p0, P, A);
}

// condense_at for a runtime layer index: the layer's values go through selects.
template <int L, bool MIX>
__device__ double condense_k(int k, const double* mu, double* th, double* q, const double* pib, const CP& P,
                             const AP& A) {
  double pt = 0, mk = 0, tk = 0, qk = 0, pk = 0;
#pragma unroll
  for (int j = 0; j < L; j++) {
    if (j < k) pt += P.grav * mu[j];
    if (j == k) { mk = mu[j]; tk = th[j]; qk = q[j]; pk = pib[j]; }
  }
  const double dq = condense_at<L, MIX>(k, tk, qk, pk, pt, mk, P, A);
#pragma unroll
  for (int j = 0; j < L; j++)
    if (j == k) { th[j] = tk; q[j] = qk; }
  return dq;
}

// The temperature at layer k's lower edge (pressure pb) (air.cu:476-482): its theta there, or for a spread middle
// layer (sat_prof) its lower theta*.
template <int L, bool MIX>
__device__ __forceinline__ double base_T(int k, double th, double pb, const CP& P, const AP& A) {
  double ea, eb;
  return (spread_layer<L>(k, th, P, A, ea, eb) ? ea : th) * fpow_<MIX>(pb / P.p0, P.kappa);
}

// The vapor of layer k's air at its top edge (top) or bottom edge (air.cu:361-378): its relative humidity against
// the layer's capacity held uniform through the layer (a spread layer's edge at its theta*); layers above edge_k (with
// L >= 3) and, with bl_q, the mixed bottom layer keep their vapor even. Air of mass m leaving through that edge takes
// the edge's vapor plus m's share of the way to the mean, at most the layer's whole vapor. k may be a runtime index.
template <int L, bool MIX>
__device__ double edge_q(int k, bool top, const double* mu, const double* th, const double* q, double m, const CP& P,
                         const AP& A) {
  using F = std::conditional_t<MIX, float, double>;
  double mk = 0, tk = 0, qk = 0, pt = 0;
#pragma unroll
  for (int j = 0; j < L; j++) {
    if (j < k) pt += P.grav * mu[j];
    if (j == k) { mk = mu[j]; tk = th[j]; qk = q[j]; }
  }
  if (!(qk > 0) || (k < (int)A.edge_k && L >= 3) || (A.bl_q > 0 && k == L - 1 && k > 0)) return qk;
  F qs, dqs;
  layer_cap_k<L, MIX>(k, tk, pt, pt + P.grav * mk, P, A, qs, dqs);
  if (!(qs > 0)) return qk;
  const double p = top ? pt : pt + P.grav * mk;
  if (!(p > 0)) return qk;
  double ea, eb;
  const bool sp = spread_layer<L>(k, tk, P, A, ea, eb);
  const double te = sp ? (top ? eb : ea) : tk;
  const double T = te * fpow_<MIX>(p / P.p0, P.kappa);
  const double qe = T < 160 ? 0.0 : qk * (double)qsat_a<F>((F)T, (F)p, (F)A.eps, A.ice_air) / qs;
  const double f = m > 0 ? fmin(1.0, m / mk) : 0.0;
  return m > 0 ? fmin(qe + (qk - qe) * f, mk * qk / m) : qe;
}

// A parcel (th, q) lifted to pressure p, condensing what it cannot hold (air.cu:468-473).
template <bool MIX>
__device__ __forceinline__ void parcel(double th, double q, double p, const CP& P, const AP& A, double& thp,
                                       double& qp) {
  thp = th;
  qp = q;
  const double ex = fpow_<MIX>(p / P.p0, P.kappa);
  condense<MIX>(thp, qp, P.cp * ex, p, P, A);
}

// The column's energy (air.cu:484-488): enthalpy, and for moist air the vapor's latent energy.
template <int L, bool MOIST = false>
__device__ __forceinline__ double column_energy(const double* mu, const double* pib, const double* th, const double* q,
                                                const CP& P) {
  double E = 0;
#pragma unroll
  for (int k = 0; k < L; k++) {
    if constexpr (MOIST) E += mu[k] * pib[k] * th[k] + P.Lv * mu[k] * q[k];
    else E += mu[k] * pib[k] * th[k];   // (dry: the vapor's latent term is zero)
  }
  return E;
}

// Move m (> 0) kg/m^2 of air from layer `from` to `to` (air.cu:495-520): the layers' masses, Exner functions, theta
// (and vapor: it leaves with q_leave per kg and arrives with q_arrive) re-divided, the column's energy residual booked
// into the receiver. CL: the air takes its cloud water cl along (the receiver's share re-divided).
template <int L, bool MOIST = false, bool CL = false>
__device__ void move_mass(const CP& P, double* mu, double* th, double* q, double* pib, int from, int to, double m,
                          double th_leave, double th_arrive, double q_leave = 0, double q_arrive = 0,
                          double* cl = nullptr) {
  if (m == 0) return;
  double pb2[L], pib2[L], mu2[L], th2[L], q2[L];
  const double Eb = column_energy<L, MOIST>(mu, pib, th, q, P);
#pragma unroll
  for (int k = 0; k < L; k++) {
    mu2[k] = mu[k];
    th2[k] = th[k];
    if constexpr (MOIST) q2[k] = q[k];
  }
  // (from and to are runtime indices; the arrays stay in registers through compile-time selects)
  double mf = 0, mt = 0, tf = 0, tt = 0, qf = 0, qt = 0;
#pragma unroll
  for (int k = 0; k < L; k++) {
    if (k == from) { mf = mu[k]; tf = th[k]; mu2[k] -= m; if constexpr (MOIST) qf = q[k]; }
    if (k == to) { mt = mu[k]; tt = th[k]; mu2[k] += m; if constexpr (MOIST) qt = q[k]; }
  }
  if (P.xf) {
    const int d = m > 0 ? from : to, r = m > 0 ? to : from;
    const double am = fabs(m);
    double mur = 0, mur2 = 0;
#pragma unroll
    for (int k = 0; k < L; k++)
      if (k == r) {
        mur = mu[k];
        mur2 = mu2[k];
      }
    for (int g = 0; g < 3; g++) P.xf[g * L + r] = (mur * P.xf[g * L + r] + am * P.xf[g * L + d]) / mur2;
  }
  layer_exner<L>(mu2, P, pb2, pib2);
  double muf2 = 0, mut2 = 0;
#pragma unroll
  for (int k = 0; k < L; k++) {
    if (k == from) muf2 = mu2[k];
    if (k == to) mut2 = mu2[k];
  }
  const double thf = (mf * tf - m * th_leave) / muf2, tht = (mt * tt + m * th_arrive) / mut2;
  double qfn = 0, qtn = 0;
  if constexpr (MOIST) {
    qfn = (mf * qf - m * q_leave) / muf2;
    qtn = (mt * qt + m * q_arrive) / mut2;
  }
#pragma unroll
  for (int k = 0; k < L; k++) {
    if (k == from) { th2[k] = thf; if constexpr (MOIST) q2[k] = qfn; }
    if (k == to) { th2[k] = tht; if constexpr (MOIST) q2[k] = qtn; }
  }
  if constexpr (CL) {
    double clf = 0, clt = 0;
#pragma unroll
    for (int k = 0; k < L; k++) {
      if (k == from) clf = cl[k];
      if (k == to) clt = cl[k];
    }
    const double cln = (mt * clt + m * clf) / mut2;
#pragma unroll
    for (int k = 0; k < L; k++)
      if (k == to) cl[k] = cln;
  }
  const double Ea = column_energy<L, MOIST>(mu2, pib2, th2, q2, P);
#pragma unroll
  for (int k = 0; k < L; k++)
    if (k == to) th2[k] += (Eb - Ea) / (mu2[k] * pib2[k]);
#pragma unroll
  for (int k = 0; k < L; k++) {
    th[k] = th2[k];
    mu[k] = mu2[k];
    pib[k] = pib2[k];
    if constexpr (MOIST) q[k] = q2[k];
  }
}

// The dry crossing of the isentropic interfaces by the heating dth (air.cu:744-850); wxo[j] the mass moved up
// across interface j. Dry air: the edge-vapor and freeze-drying options move the same mass without vapor.
// et[k], eb[k]: the heating at layer k's top and bottom edge (the reference's dtop[k] or dth[k], and dbot[k] or dth[k],
// the top layer's bottom-edge default its dth_up): always given, so every array stays in registers. CL: the crossing
// air takes its cloud water cl along.
template <int L, bool MOIST = false, bool CL = false, bool MIX = false>
__device__ void crossing(const CP& P, double* mu, double* th, double* q, const double* dth, const double* et,
                         const double* eb, double* pib, double* wxo, double* cl = nullptr, const AP* A = nullptr,
                         double* gross = nullptr, size_t gstride = 0) {
#pragma unroll
  for (int j = 0; j + 1 < L; j++) wxo[j] = 0;
  if (P.theta_star <= 0) return;
  double dmj[L], thcj[L], efa[L], efb[L];
  bool got[L];
#pragma unroll
  for (int k = 0; k < L; k++) got[k] = false;
  // a layer's edge densities, from its theta before the heating, once per layer (only layers below the top one: the
  // top layer has no upper interface, and no branch asks for it)
  auto dens = [&](int k, double& fa, double& fb) {
#pragma unroll
    for (int kk = 1; kk < L; kk++)
      if (kk == k) {
        if (!got[kk]) {
          edge_density(tstar<L>(kk, P), tstar<L>(kk - 1, P), th[kk] - dth[kk], efa[kk], efb[kk]);
          got[kk] = true;
        }
        fa = efa[kk];
        fb = efb[kk];
      }
  };
  const bool blm = P.bl_mix > 0;
#pragma unroll
  for (int j = 0; j + 1 < L; j++) {
    const int a = j, b = j + 1;
    const double ts_ = tstar<L>(j, P);
    double dm = 0, thc = ts_, thd = ts_, gdown = 0;   // (gdown: the gross down part, for the diagnostics rows)
    if (dth[b] > 0 && blm && b == L - 1) {
      if (th[b] >= ts_) {
        double fa, fb;
        dens(a, fa, fb);
        dm = -fmin(fmin(mu[a] * dth[b] * fa, fmax(mu[a] - P.mu_bl, 0.0) * P.bl_cap), P.bl_room);
        thd = ts_ - 0.5 * dth[b];
      }
    } else if (blm && b < L - 1 && th[b] >= ts_) {
      dm = mu[b] - P.floor_;
      thc = th[b];
    } else if (et[b] > 0) {
      const double eh = et[b];
      if (th[b] >= ts_) {
        dm = mu[b] - P.floor_;
        thc = th[b];
      } else if (b == L - 1) {
        dm = mu[b] * eh / (2 * (ts_ - (th[b] - dth[b])));
      } else {
        double fa, fb;
        dens(b, fa, fb);
        dm = mu[b] * eh * fb;
      }
    }
    const double da = eb[a];
    if (blm && a >= 1 && th[a] <= ts_ && dm <= 0) {
      const double down = mu[a] - P.floor_;
      thd = th[a];
      dm -= down;
      gdown = down;
    } else if (da < 0) {
      double down;
      if (th[a] <= ts_) {
        down = mu[a] - P.floor_;
        thd = th[a];
      } else if (a == 0) {
        down = mu[a] * (-da) / (2 * ((th[a] - dth[a]) - ts_));
      } else {
        double fa, fb;
        dens(a, fa, fb);
        down = mu[a] * (-da) * fa;
      }
      dm -= down;
      gdown = down;
    }
    if (gross) {
      gross[(size_t)(2 * j) * gstride] = dm + gdown;
      gross[(size_t)(2 * j + 1) * gstride] = gdown;
    }
    dmj[j] = dm;
    thcj[j] = dm > 0 ? thc : thd;
  }
#pragma unroll
  for (int j = L - 2; j >= 0; j--) {
    const int a = j, b = j + 1;
    double dm = dmj[j];
    const double maxup = mu[b] - P.floor_, maxdn = mu[a] - P.floor_;
    if (dm > maxup) dm = maxup > 0 ? maxup : 0.0;
    if (-dm > maxdn) dm = maxdn > 0 ? -maxdn : 0.0;
    // (the crossing air carries its layer's vapor)
    double qb_ = 0, qa_ = 0;
    if constexpr (MOIST) {
#pragma unroll
      for (int k = 0; k < L; k++) {
        if (k == b) qb_ = q[k];
        if (k == a) qa_ = q[k];
      }
    }
    // edge vapor: the crossing air carries the vapor at the edge it leaves through; freeze-drying: air rising into
    // the stratosphere past the cold point at theta* holds no more than saturation there, the rest staying below with
    // the layers' cloud water
    const bool ev = MOIST && A && A->edge_vap > 0;
    if (MOIST && A && dm > 0 && a == (int)A->edge_k - 1 && L >= 3 && A->freeze_dry > 0) {
      using F = std::conditional_t<MIX, float, double>;
      double pi_ = 0;
#pragma unroll
      for (int k = 0; k <= a; k++) pi_ += P.grav * mu[k];
      const double Ti = tstar<L>(a, P) * fpow_<MIX>(pi_ / P.p0, P.kappa);
      const double qx = fmin(ev ? edge_q<L, MIX>(b, true, mu, th, q, dm, P, *A) : qb_,
                             Ti < 160 ? 0.0 : (double)qsat_a<F>((F)Ti, (F)pi_, (F)A->eps, A->ice_air));
      const double ma = mu[a], mb = mu[b], ca = CL ? cl[a] : 0.0, cb = CL ? cl[b] : 0.0;
      move_mass<L, MOIST, CL>(P, mu, th, q, pib, b, a, dm, thcj[j], thcj[j], qx, qx, cl);
      if constexpr (CL) {
        cl[a] = ma * ca / mu[a];
        cl[b] = mb * cb / mu[b];
      }
    } else if (dm > 0) {
      const double qx = ev ? edge_q<L, MIX>(b, true, mu, th, q, dm, P, *A) : qb_;
      move_mass<L, MOIST, CL>(P, mu, th, q, pib, b, a, dm, thcj[j], thcj[j], qx, qx, cl);
    } else if (dm < 0) {
      const double qx = ev ? edge_q<L, MIX>(a, false, mu, th, q, -dm, P, *A) : qa_;
      move_mass<L, MOIST, CL>(P, mu, th, q, pib, a, b, -dm, thcj[j], thcj[j], qx, qx, cl);
    }
    wxo[j] = dm;
  }
  const double tiny = P.floor_ * 1.000001;
#pragma unroll
  for (int j = L - 2; j >= 0; j--) {
    const int a = j, b = j + 1;
    const double ts_ = tstar<L>(j, P);
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

// hs_column (air.cu:2065-2109): each layer relaxed toward Held & Suarez's equilibrium theta at its mass-mean pressure,
// then the interfaces crossed; theta updated in place, wx (L-1, N) the mass to move up, qnet the heating (W/m^2),
// stored the column's enthalpy change (J/m^2). The layer masses are left to the exchange that moves wx.
template <int L>
__global__ void __launch_bounds__(BS) k_hs_column(int N, CP P, double ka, double ks, const double* __restrict__ lat,
                                                  const double* __restrict__ mu_in, double* __restrict__ theta,
                                                  double dt, double* __restrict__ wx, double* __restrict__ qnet,
                                                  double* __restrict__ stored) {
  const int i = XTID;
  if (i >= N) return;
  double mu[L], th[L], pb[L], pib[L], dth[L], wxo[L];
#pragma unroll
  for (int k = 0; k < L; k++) {
    mu[k] = mu_in[(size_t)k * N + i];
    th[k] = theta[(size_t)k * N + i];
  }
  layer_exner<L>(mu, P, pb, pib);
  double E0 = 0, Qs = 0;
#pragma unroll
  for (int k = 0; k < L; k++) E0 += mu[k] * pib[k] * th[k];
  const double s2 = sin(lat[i]) * sin(lat[i]), c2 = 1 - s2;
  const double ps = pb[L - 1];
  double pt = 0;
#pragma unroll
  for (int k = 0; k < L; k++) {
    const double pbar = pow(pib[k] / P.cp, 1 / P.kappa) * P.p0;
    double Teq = (315 - 60 * s2 - 10 * log(pbar / P.p0) * c2) * pow(pbar / P.p0, P.kappa);
    if (Teq < 200) Teq = 200;
    const double theq = Teq * P.cp / pib[k];
    const double st = pt / ps, sb = pb[k] / ps;
    const double lo = st > 0.7 ? st : 0.7;
    const double f = sb > 0.7 ? ((sb - 0.7) * (sb - 0.7) - (lo - 0.7) * (lo - 0.7)) / (2 * 0.3 * (sb - st)) : 0.0;
    const double kt = ka + (ks - ka) * f * c2 * c2;
    const double d = (theq - th[k]) * (1 - exp(-kt * dt));
    dth[k] = d;
    th[k] += d;
    Qs += mu[k] * pib[k] * d / dt;
    pt = pb[k];
  }
  double eb[L];
#pragma unroll
  for (int k = 0; k < L; k++) eb[k] = dth[k];
  double dth_up = dth[0];
  if (L == 2) {
    const double pi_ = pb[0], si = pi_ / ps;
    double Ti = (315 - 60 * s2 - 10 * log(pi_ / P.p0) * c2) * pow(pi_ / P.p0, P.kappa);
    if (Ti < 200) Ti = 200;
    const double thi = Ti * pow(P.p0 / pi_, P.kappa);
    const double ki = ka + (ks - ka) * (si > 0.7 ? (si - 0.7) / 0.3 : 0.0) * c2 * c2;
    dth_up = (thi - P.theta_star) * (1 - exp(-ki * dt));
  }
  eb[0] = dth_up;
  crossing<L>(P, mu, th, nullptr, dth, dth, eb, pib, wxo);
#pragma unroll
  for (int j = 0; j + 1 < L; j++) wx[(size_t)j * N + i] = wxo[j];
  double E1 = 0;
#pragma unroll
  for (int k = 0; k < L; k++) {
    theta[(size_t)k * N + i] = th[k];
    E1 += mu[k] * pib[k] * th[k];
  }
  qnet[i] = Qs;
  stored[i] = E1 - E0;
}

#if !ISW_COLUMN_CLOUDS
// surface_layer (air.cu:2015-2061): rho C_D U and rho C_H U by Louis's bulk form over the sea (Charnock), sea ice and
// land, four fixed-point passes with the free convection's gusts.
__global__ void __launch_bounds__(BS) k_surface_layer(int N, int L, CP P, double rd, double vf, double eps,
                                                      const double* __restrict__ uabs, const double* __restrict__ mu_in,
                                                      const double* __restrict__ theta, const double* __restrict__ qv,
                                                      const double* __restrict__ Ts, const double* __restrict__ ice,
                                                      const double* __restrict__ land, double* __restrict__ cd,
                                                      double* __restrict__ ch, const double* __restrict__ lc) {
  const int i = XTID;
  if (i >= N) return;
  const int bot = L - 1;
  double ps = 0;
  for (int k = 0; k < L; k++) ps += P.grav * mu_in[(size_t)k * N + i];
  const double mb = mu_in[(size_t)bot * N + i], th = theta[(size_t)bot * N + i];
  const double Ta = th * pow(ps / P.p0, P.kappa), T0 = Ts[i];
  const double rho = ps / (rd * Ta);
  const double qa = qv ? qv[(size_t)bot * N + i] : 0.0;
  const bool isl = land && land[i] > 0.5;
  const double q0 = qv && P.moist > 0 && !isl
                        ? (P.q_sea > 0 ? P.q_sea : 1.0) * (eps * 611.0 * exp(17.27 * (T0 - 273.15) / (T0 - 35.85)) / ps)
                        : qa;
  const double fice = qv && P.moist > 0 && ice && !isl ? fmin(1.0, ice[i] / P.rho_ice / P.ice_cover) : 0.0;
  const double depth = rd * Ta / P.grav * log(ps / fmax(ps - P.grav * mb, 1.0));
  const double zr = fmax(10.0, 0.1 * depth);
  const double dthv = (Ta - T0) + vf * Ta * (qa - q0);
  const double K = 0.4, nu = 1.5e-5, b = 5, c = 5;
  const double u2 = uabs[i] * uabs[i];
  double U = sqrt(u2 + P.bulk_gust * P.bulk_gust), Cd = 1e-3, Ch = 1e-3;
  for (int it = 0; it < 4; it++) {
    const double us = fmax(sqrt(Cd) * U, 1e-3);
    double z0, zh;
    if (isl) {
      z0 = lc ? lc[i] : P.z0_land;
      zh = 0.1 * z0;
    } else {
      const double zs = 0.011 * us * us / P.grav + 0.11 * nu / us, zsh = 0.40 * nu / us;
      z0 = (1 - fice) * zs + fice * 1e-3;
      zh = (1 - fice) * zsh + fice * 1e-4;
    }
    const double lm = log(zr / z0), lh = log(zr / zh);
    const double CdN = K * K / (lm * lm), ChN = K * K / (lm * lh);
    const double Ri = P.grav * zr * dthv / (Ta * U * U);
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
    const double ws = B > 0 ? cbrt(P.grav / Ta * B * depth) : 0.0;
    U = sqrt(u2 + P.bulk_gust * P.bulk_gust + 1.44 * ws * ws);
  }
  cd[i] = rho * Cd * U;
  ch[i] = rho * Ch * U;
}

// cell_speed (air.cu:1978-1989): the bottom layer's wind speed at cells from its edges' kinetic energy.
__global__ void __launch_bounds__(BS) k_cell_speed(int N, const double* __restrict__ u, const int* __restrict__ cellEdge,
                                                   const double* __restrict__ le, const double* __restrict__ de,
                                                   const double* __restrict__ areaCell, double* __restrict__ out) {
  const int i = XTID;
  if (i >= N) return;
  const int2* ce = reinterpret_cast<const int2*>(cellEdge + (size_t)i * 6);
  const int2 p0 = ce[0], p1 = ce[1], p2 = ce[2];
  const int e[6] = {p0.x, p0.y, p1.x, p1.y, p2.x, p2.y};
  double ke = 0;
#pragma unroll
  for (int m = 0; m < 6; m++)
    if (e[m] >= 0) ke += le[e[m]] * de[e[m]] * 0.25 * u[e[m]] * u[e[m]];
  out[i] = sqrt(2 * ke / areaCell[i]);
}

// cell_speed_perot (air.cu:1994-2006): the speed of Perot's reconstructed cell wind, op (N, 6, 2) east and north
// coefficients per edge slot.
__global__ void __launch_bounds__(BS) k_cell_speed_perot(int N, const double* __restrict__ u,
                                                         const int* __restrict__ cellEdge,
                                                         const double* __restrict__ op, double* __restrict__ out) {
  const int i = XTID;
  if (i >= N) return;
  const int2* ce = reinterpret_cast<const int2*>(cellEdge + (size_t)i * 6);
  const int2 p0 = ce[0], p1 = ce[1], p2 = ce[2];
  const int e[6] = {p0.x, p0.y, p1.x, p1.y, p2.x, p2.y};
  double ue = 0, un = 0;
#pragma unroll
  for (int m = 0; m < 6; m++)
    if (e[m] >= 0) {
      ue += op[((size_t)i * 6 + m) * 2] * u[e[m]];
      un += op[((size_t)i * 6 + m) * 2 + 1] * u[e[m]];
    }
  out[i] = sqrt(ue * ue + un * un);
}

// gas_src (air.cu:2133-2137): each gas's surface source (mole fraction x kg/m^2/s) into the bottom layer over dt.
__global__ void __launch_bounds__(BS) k_gas_src(int N, int L, int G, const double* __restrict__ src,
                                                const double* __restrict__ h, double dt, double* __restrict__ x) {
  const int i = XTID;
  if (i >= N) return;
  const double hb = h[(size_t)(L - 1) * N + i];
  for (int g = 0; g < G; g++) x[((size_t)g * L + L - 1) * N + i] += src[(size_t)g * N + i] * dt / hb;
}

// advance_time (air.cu:1975): the device clock, advanced inside captured ticks.
__global__ void k_advance_time(double* t, double dt) { t[0] += dt; }
#endif  // !ISW_COLUMN_CLOUDS

// ---------------------------------------------------------------------------------------------------------------
// air_column (air.cu:876-1972): the dry column, and the moist one over the sea and land, with clouds and the mixed
// boundary layer: the sun, the shortwave's split, the gray longwave (the vapor's part from the column's vapor with
// lw_w0) or the k-tables' fluxes, the slab surface, the sea-ice skin or the ground under the bulk or surface-layer
// exchange, the evaporation, the heating as theta, the condensation (the cloud scheme), the crossing with the
// k-tables' edge heating, the mixed layer's moves, and the moist convection's parcel plume. MIX selects the reference's mixed build (air.cu:24-36): 1 - e^-x, the edge Exner powers and the
// saturation and condensation arithmetic (its `fast` type) in float32; the rest is FP64 in both forms.
// ---------------------------------------------------------------------------------------------------------------

// The cloud statistics (air.cu:383-388): total water qt spread evenly over +-(1 - rh_crit) qs about its mean; the
// saturated share f and the mean condensate l there.
__device__ __forceinline__ void cloud_pdf(double qt, double qs, const AP& A, double& l, double& f) {
  const double b = (1 - A.rh_crit) * qs, s = qt - qs;
  if (s <= -b) {
    l = 0;
    f = 0;
  } else if (s >= b) {
    l = s;
    f = 1;
  } else {
    f = (s + b) / (2 * b);
    l = (s + b) * f / 2;
  }
}

// The cover of cloud water cw spread at a cloud's typical water content by phase (air.cu:393-396).
__device__ __forceinline__ double cover_of(double cw, double mu, double liq, const AP& A, double qs) {
  const double s = A.cw_qs > 0 && qs > 0 ? fmin(1.0, A.cw_qs * qs / A.q_auto) : 1.0;
  return fmin(1.0, cw / (s * (liq * A.q_auto + (1 - liq) * A.l_ice)) * fmax(1.0, mu / A.mu_cloud));
}

// The mixed bottom layer's cloud statistics (air.cu:397-426, bl_q): its vapor even through its depth, it is nearest
// saturation at its top: at each of the capacity's four levels, l the mean condensate, f the cover where it is most,
// dqf the condensate's fall per kelvin of warming; ml_clear (unless pdf): only air actually saturated condenses.
// False for any other layer (or without bl_q).
template <int L, bool MIX>
__device__ bool mixed_cloud(int k, const double* mu, double th, double qt, const CP& P, const AP& A, double& l,
                            double& f, double& dqf, bool pdf) {
  using F = std::conditional_t<MIX, float, double>;
  if (!(A.bl_q > 0 && k == L - 1 && k > 0)) return false;
  constexpr double x[4] = {-0.8611363115940526, -0.3399810435848563, 0.3399810435848563, 0.8611363115940526};
  constexpr double wq[4] = {0.3478548451374538, 0.6521451548625461, 0.6521451548625461, 0.3478548451374538};
  double pt = 0;
#pragma unroll
  for (int j = 0; j < L - 1; j++) pt += P.grav * mu[j];
  const double pb = pt + P.grav * mu[L - 1];
  l = 0;
  f = 0;
  dqf = 0;
#pragma unroll
  for (int j = 0; j < 4; j++) {
    const double p = 0.5 * (pt + pb) + 0.5 * (pb - pt) * x[j];
    const double ex = fpow_<MIX>(p / P.p0, P.kappa), T = th * ex;
    if (T < 160) continue;
    const double sq = (double)qsat_a<F>((F)T, (F)p, (F)A.eps, A.ice_air);
    double lj, fj;
    if (A.ml_clear > 0 && !pdf) {
      lj = fmax(0.0, qt - sq);
      fj = qt > sq ? 1.0 : 0.0;
    } else {
      cloud_pdf(qt, sq, A, lj, fj);
    }
    l += 0.5 * wq[j] * lj;
    f = fmax(f, fj);
    dqf += 0.5 * wq[j] * fj * sq * P.Lv / (A.Rv * T * T) * ex;
  }
  return true;
}

// Cloud water in layer k (a runtime index) condensing up to the cloud statistics' condensate with the latent heat's
// feedback on the capacity, or eroding over tau_erode (air.cu:438-450); returns the condensate.
template <int L, bool MIX>
__device__ double cloud_adjust_k(int k, const double* mu, double* th, double* q, double* cl, const double* pib,
                                 double dt, const CP& P, const AP& A) {
  using F = std::conditional_t<MIX, float, double>;
  double pt = 0, mk = 0, tk = 0, qk = 0, ck = 0, pk = 0;
#pragma unroll
  for (int j = 0; j < L; j++) {
    if (j < k) pt += P.grav * mu[j];
    if (j == k) { mk = mu[j]; tk = th[j]; qk = q[j]; ck = cl[j]; pk = pib[j]; }
  }
  double l, f, dqa, dqm;
  if (mixed_cloud<L, MIX>(k, mu, tk, qk + ck, P, A, l, f, dqm, false)) {
    f = 1;
    dqa = dqm;
  } else {
    F qs, dqs;
    layer_cap_k<L, MIX>(k, tk, pt, pt + P.grav * mk, P, A, qs, dqs);
    if (!(qs > 0)) return 0;
    cloud_pdf(qk + ck, qs, A, l, f);
    dqa = dqs;
  }
  const double dl = l > ck ? (l - ck) / (1 + P.Lv / pk * dqa * f) : -(ck - l) * omexp<MIX>(dt / A.tau_erode);
#pragma unroll
  for (int j = 0; j < L; j++)
    if (j == k) {
      cl[j] = ck + dl;
      q[j] = qk - dl;
      th[j] = tk + P.Lv * dl / pk;
    }
  return dl;
}

// The liquid share of condensate at T (air.cu:160): 1 at freezing, 0 at -20 C, in the build's `fast` type.
template <bool MIX>
__device__ __forceinline__ double liq_of(double T) {
  if constexpr (MIX) return fminf(1.0f, fmaxf(0.0f, ((float)T - 253.15f) / 20.0f));
  else return fmin(1.0, fmax(0.0, (T - 253.15) / 20.0));
}

// The deep plume's full ascent (air.cu:540-666; its options plume_ent, ent_cloud, ent_det, ice_air, ent_ifs, ent_loc,
// det_all, buoy_v): bottom air (th0,
// q0) lifted from the bottom layer's top (pi_) up through the layers in M sub-levels each, condensing as it goes,
// mixing with (plume_ent) and, through the layer above the bottom one, entraining (ent_cloud: its mass growing as
// exp(ent_cloud z); ent_det: as much detrained) the air it passes, whose theta at mass share x above the layer's
// lower edge a is a + (b - a) ln(1 + x (e^s - 1)) / s. A: its buoyancy summed by mass up to the highest level it beats
// (the bottom layer's stretch a triangle); kd, thd, qd: the layer it detrains into, its theta and vapor there; qx: what
// its mixing gave each layer per kg; G: its mass per kg leaving the bottom layer; FZ: its frozen condensate; lt: its
// condensate freezing between 0 and -40 C, those levels' pressures and its CAPE (for lightning). ent_ifs: the
// entraining layers' entrainment and detrainment the IFS's, 1.75e-3 /m (1.3 - RH) (q_s / q_s at cloud base)^3 and
// 0.75e-4 /m, in sub-levels of ~25 hPa (air.cu:578-584, 609-624); det_all: every tropospheric layer (below edge_k's
// stratosphere) entraining; ent_loc: the air taken in holding the layer's vapor at its relative humidity level by
// level (air.cu:586-598); buoy_v: buoyancy in virtual theta, theta (1 + 0.608 q) (air.cu:564-567); cu_dth (the
// cumulus layer, air.cu:580-583, 609-611, 654): the cloud layer above the cumulus layer entraining too, each entraining
// layer's first sub-level taking the rest of the last one below, the shallow plume going on into that cloud layer; Gl
// (L): what it took in from each layer with ent_ifs, per kg leaving the bottom layer. The layer loop is unrolled with a
// stop flag, so every per-layer value stays in registers.
template <int L, bool MIX>
__device__ void plume_ascent(const double* mu, const double* th, const double* q, double th0, double q0, double pi_,
                             const CP& P, const AP& A, double& Ab, int& kd, double& thd, double& qd, double* qx,
                             double* lt, double& G, double& FZ, double* Gl, bool shal = false) {
  constexpr int bot = L - 1, M = bot > 2 ? 3 : 8;
  const bool cu = A.cu_dth > 0;
  double Mp = 1, tp = th0, qp = q0, fz = 0, gk[L], dzc = 0;
  auto lift = [&](double p_) {
    const double q_ = qp, ex_ = fpow_<MIX>(p_ / P.p0, P.kappa);
    condense<MIX>(tp, qp, P.cp * ex_, p_, P, A);
    if (A.ice_air > 0 && q_ > qp) {
      const double z_ = (q_ - qp) * (1 - liq_of<MIX>(tp * ex_));
      tp += A.Lf * z_ / (P.cp * ex_);
      fz += Mp * z_;
    }
    return ex_;
  };
  lift(pi_);
  auto vf = [&](double qv_) { return A.buoy_v > 0 ? 1 + 0.608 * qv_ : 1.0; };
  const bool ifs = A.ent_ifs > 0;
  double sum = 0.5 * mu[bot] * (tp * vf(qp) - tstar<L>(bot - 1, P) * vf(q[bot - 1])), pt = pi_;
  // (ent_ifs: the environment's saturation at cloud base, the mixed layer's top)
  using F = std::conditional_t<MIX, float, double>;
  auto qsat_ = [&](double T, double p) { return (double)qsat_a<F>((F)T, (F)p, (F)A.eps, A.ice_air); };
  const double qsb = ifs ? qsat_(th[bot] * fpow_<MIX>(pi_ / P.p0, P.kappa), pi_) : 1.0;
  Ab = 0;
  kd = -1;
  thd = tp;
  qd = qp;
  double cape = 0, q0c = -1, q40 = -1, p0c = 0, p40 = 0, qa[L];
#pragma unroll
  for (int k = 0; k < L; k++) {
    qa[k] = 0;
    qx[k] = 0;
    gk[k] = 0;
    Gl[k] = 0;
  }
  bool stop = false;
#pragma unroll
  for (int k = bot - 1; k >= 0; k--) {
    if (stop) continue;
    const double a = tstar<L>(k, P), b = k > 0 ? tstar<L>(k - 1, P) : 0.0;
    const double s = k > 0 ? edge_shape(a, b, th[k]) : 0.0, es = k > 0 && !MIX ? expm1(s) : 0.0;
    const bool ek = k > 0 && (k == bot - 1 || (cu && ifs && k == bot - 2) ||
                              (A.det_all > 0 && ifs && !shal && k >= (int)A.edge_k));
    const int Mk = ifs && ek ? min(48, max(M, (int)ceil(P.grav * mu[k] / 2500.0))) : M;
    const double f = A.plume_ent > 0 ? omexp<MIX>(A.plume_ent * P.grav * mu[k] / Mk) : 0.0;
    const bool loc = A.ent_loc > 0 && A.ent_cloud > 0 && ifs && ek && k >= (int)A.edge_k && q[k] > 0;
    double qsl = 0;
    if (loc)
      for (int j = 0; j < Mk; j++) {
        const double x = (j + 0.5) / Mk, pl = pt - x * P.grav * mu[k];
        const double te = a + (b - a) * spread_of<MIX>(x, s, es);
        qsl += qsat_(te * fpow_<MIX>(pl / P.p0, P.kappa), pl) / Mk;
      }
    bool last = false;
    for (int j = 0; j < Mk; j++) {
      const double x = (j + 0.5) / Mk, pl = pt - x * P.grav * mu[k];
      const double ex = lift(pl);
      const double te = k == 0 ? th[0] : a + (b - a) * spread_of<MIX>(x, s, es);
      if (A.ent_cloud > 0 && ek) {
        double dz = A.rd * te * ex / P.grav * (P.grav * mu[k] / Mk) / pl;
        if (ifs && j == 0) dz *= 0.5;                  // (from cloud base to the first sub-level's middle)
        if (cu && j == 0) dz += dzc;                   // (and the rest of the layer below's last sub-level)
        if (cu && j == Mk - 1) dzc = 0.5 * A.rd * te * ex / P.grav * (P.grav * mu[k] / Mk) / pl;
        if (ifs) {
          const double qse = qsat_(te * ex, pl), qen = loc && qsl > 0 ? q[k] * qse / qsl : q[k];
          const double rh = fmin(1.0, fmax(0.0, qen / qse)), fs = fmin(1.0, qse / qsb);
          const double eps_ = (shal ? 3.5e-3 : 1.75e-3) * (1.3 - rh) * fs * fs * fs, del_ = fmin(eps_, 0.75e-4);
          const double fe_ = -expm1(-eps_ * dz), fd_ = -expm1(-del_ * dz);
          qa[k] += Mp * fd_ * (qp - qen);
          if (loc) qa[k] -= Mp * expm1((eps_ - del_) * dz) * (qen - q[k]);   // (the growth taken in at qen)
          tp += fe_ * (te - tp);
          qp += fe_ * (qen - qp);
          gk[k] += Mp * expm1((eps_ - del_) * dz);
          Mp *= exp((eps_ - del_) * dz);
        } else {
          const double g_ = expm1(A.ent_cloud * dz), fe = g_ / (1 + g_);
          if (A.ent_det > 0) qa[k] += Mp * fe * (qp - q[k]);
          tp += fe * (te - tp);
          qp += fe * (q[k] - qp);
          if (!(A.ent_det > 0)) Mp *= 1 + g_;
        }
      }
      if (A.plume_ent > 0) {
        qa[k] += Mp * f * (qp - q[k]);
        tp += f * (te - tp);
        qp += f * (q[k] - qp);
      }
      const double tpv = tp * vf(qp), tev = te * vf(q[k]);
      sum += mu[k] / Mk * (tpv - tev);
      last = tpv > tev;
      const double Tp = tp * ex;
      if (tpv > tev) cape += A.rd * (tpv - tev) * ex * P.grav * mu[k] / Mk / pl;
      if (q0c < 0 && Tp < A.Tf) {
        q0c = qp;
        p0c = pl;
      }
      if (q40 < 0 && Tp < A.Tf - 40) {
        q40 = qp;
        p40 = pl;
      }
      if (last) {
        Ab = sum;
        kd = k;
        thd = tp;
        qd = qp;
        G = Mp;
        FZ = fz;
#pragma unroll
        for (int m = 0; m < L; m++) {
          qx[m] = qa[m];
          Gl[m] = gk[m];
        }
      }
    }
    // (the shallow plume stops in the cloud layer; with cu_dth in the cumulus layer or the cloud layer above it)
    if (!last || k == 0 || (shal && !(cu && k == bot - 1))) {
      stop = true;
      continue;
    }
    pt -= P.grav * mu[k];
    lift(pt);
    if (!(tp * vf(qp) > b * vf(k > 0 ? q[k - 1] : 0.0))) stop = true;
  }
  if (q0c >= 0 && q40 < 0) {
    q40 = qd;
    p40 = pt;
  }
  lt[0] = q0c >= 0 ? fmax(0.0, q0c - q40) : 0.0;
  lt[1] = q0c >= 0 ? p0c : 0.0;
  lt[2] = q0c >= 0 ? p40 : 0.0;
  lt[3] = cape;
}

// Lightning from one storm's physics (air.cu:718-729; Deierling et al. 2008): 9.0e-15 f_p f_np flashes per minute,
// f_p the falling graupel's mass flux and f_np the rising small ice's, the plume's ice flux shared among storms of
// updraft area storm_area rising at half sqrt(2 CAPE); only an updraft faster than graupel falls holds it up to
// charge. Flashes per second per m^2.
template <bool MIX>
__device__ double storm_flashes(double mflux, const double* lt, const CP& P, const AP& A) {
  const double qi = lt[0], w = 0.5 * sqrt(2 * fmax(0.0, lt[3])), gv = A.graupel_v;
  if (!(mflux > 0) || !(qi > 0) || !(w > gv)) return 0;
  const double rho = lt[1] / (A.rd * A.Tf), phi = rho * w * A.storm_area, lift = 1 - gv * gv / (w * w);
  double lg;
  if constexpr (MIX) lg = (double)logf((float)(lt[1] / lt[2]));
  else lg = log(lt[1] / lt[2]);
  const double H = A.rd * (A.Tf - 20) / P.grav * lg;
  const double up = phi * qi * lift, fnp = up * A.ice_small, fp = up * (1 - A.ice_small) * H / w * gv;
  return mflux / phi * 9.0e-15 * fp * fnp / 60;
}

// Air (th, q) at pressure p evaporating water until saturated (air.cu:453-466): three Newton steps, Tetens' slope.
template <bool MIX>
__device__ void wet_bulb(double& th, double& q, double p, const CP& P, const AP& A) {
  using F = std::conditional_t<MIX, float, double>;
  const double ex = fpow_<MIX>(p / P.p0, P.kappa);
  for (int it = 0; it < 3; it++) {
    const double T = th * ex;
    if (T < 160) return;
    const double qs = (double)qsat_a<F>((F)T, (F)p, (F)A.eps, A.ice_air), d = qs - q;
    if (!(d > 0)) return;
    const double dqs = qs * 17.27 * (273.15 - 35.85) / ((T - 35.85) * (T - 35.85));
    const double e = d / (1 + P.Lv / P.cp * dqs);
    q += e;
    th -= P.Lv * e / (P.cp * ex);
  }
}

// The mixed bottom layer's mass up to the level where its air (theta th, vapor q), lifted dry-adiabatically,
// saturates (air.cu:1413-1421): 40 bisections in pressure; 0 if saturated at the ground, 1e30 if not even at a tenth of
// the surface pressure.
template <int L, bool MIX>
__device__ double lcl_mass(const double* mu, double th, double q, const CP& P, const AP& A) {
  using F = std::conditional_t<MIX, float, double>;
  double ps_ = 0;
#pragma unroll
  for (int k = 0; k < L; k++) ps_ += P.grav * mu[k];
  auto sat = [&](double p_) {
    return (double)qsat_a<F>((F)(th * fpow_<MIX>(p_ / P.p0, P.kappa)), (F)p_, (F)A.eps, A.ice_air) <= q;
  };
  if (sat(ps_)) return 0;
  if (!sat(0.1 * ps_)) return 1e30;
  double lo = 0.1 * ps_, hi = ps_;
  for (int it = 0; it < 40; it++) {
    const double m_ = 0.5 * (lo + hi);
    if (sat(m_)) lo = m_;
    else hi = m_;
  }
  return (ps_ - 0.5 * (lo + hi)) / P.grav;
}

// cu_top (air.cu:668-680): the theta the bottom layer's air reaches lifted from its top by cu_dp (Pa), condensing on
// the way, in eight equal steps (no higher than 1000 Pa).
template <int L, bool MIX>
__device__ double cu_top(const double* mu, const double* th, const double* q, const CP& P, const AP& A) {
  constexpr int bot = L - 1;
  double pb = 0;
#pragma unroll
  for (int k = 0; k < bot; k++) pb += P.grav * mu[k];
  double thp = th[bot], qp = q[bot];
  for (int s = 1; s <= 8; s++) {
    const double p = fmax(pb - A.cu_dp * s / 8, 1000.0);
    condense<MIX>(thp, qp, P.cp * fpow_<MIX>(p / P.p0, P.kappa), p, P, A);
  }
  return thp;
}

// The cumulus layer's floating edges (air.cu:913-922, 1403-1407): its top (interface L - 3) at cu_dth above theta*,
// or with cu_dp at cu_top, kept 0.5 K above theta* and below the layer's mean less 1 K and the edge above less 0.5 K;
// with cl_float the cloud layer's lower edge (interface L - 4) cl_float above cu_top where that is above its fixed
// value ths_cl, below that layer's mean less 1 K and the edge above less 0.5 K.
template <int L, bool MIX>
__device__ void cu_edges(const double* mu, const double* th, const double* q, CP& P, const AP& A, double ths_cl) {
  if constexpr (L > 4) {
    if (A.cl_float > 0 && A.cu_dth > 0 && A.cu_dp > 0)
      P.ths[L - 4] = fmax(ths_cl, fmin(cu_top<L, MIX>(mu, th, q, P, A) + A.cl_float,
                                       fmin(th[L - 4] - 1, P.ths[L - 5] - 0.5)));
  }
  if constexpr (L > 3) {
    if (A.cu_dth > 0) {
      const double top = A.cu_dp > 0 ? cu_top<L, MIX>(mu, th, q, P, A) : P.theta_star + A.cu_dth;
      P.ths[L - 3] = fmax(P.theta_star + 0.5, fmin(top, fmin(th[L - 3] - 1, P.ths[L - 4] - 0.5)));
    }
  }
}

// plume_out (air.cu:682-713): a plume of base mass m from the bottom layer (theta th0, vapor q0) detraining into layer
// kd at (thp, qc), having taken in Gl[k] per kg from each layer k and given qx[k] by mixing. m is first cut to the
// layers' air above the floor and half their vapor. The air taken in below kd rises with it; kd's own leaves at the
// layer's vapor and comes back at the plume's. cond: its condensate per kg (det of it left as cloud, at most ldet per
// kg of plume); G its mass at kd per kg; the column's energy kept, the residual into kd.
template <int L, bool CL>
__device__ void plume_out(const CP& P, double* mu, double* th, double* q, double* pib, double* cl, int kd, double& m,
                          double th0, double q0, double thp, double qc, const double* Gl, const double* qx,
                          double* wxo, double ldet, double& cond, double& det, double& G) {
  constexpr int bot = L - 1;
  m = fmin(m, mu[bot] - P.floor_);
#pragma unroll
  for (int k = 0; k < bot; k++) {
    if (Gl[k] > 0) m = fmin(m, fmax(0.0, mu[k] - P.floor_) / Gl[k]);
    const double take = Gl[k] * q[k] - qx[k];
    if (take > 0) m = fmin(m, 0.5 * mu[k] * q[k] / take);
  }
  G = 1;
  cond = 0;
  det = 0;
  if (!(m > 0)) {
    m = 0;
    return;
  }
  double qin = 0, xs = 0;
#pragma unroll
  for (int k = 0; k < bot; k++) {
    G += Gl[k];
    qin += Gl[k] * q[k];
    xs += qx[k];
  }
  const double cond_ = q0 + qin - G * qc - xs;
  cond = fmax(0.0, cond_);
  det = CL ? fmin(cond, ldet * G) : 0.0;
  const double qarr = qc + (cond_ - cond) / G;
  double tk[L], qk[L];
#pragma unroll
  for (int k = 0; k < L; k++) {
    tk[k] = th[k];
    qk[k] = q[k];
  }
  const double Ec = column_energy<L, true>(mu, pib, th, q, P);
#pragma unroll
  for (int k = 0; k < bot; k++) q[k] = fmax(0.0, q[k] + m * qx[k] / mu[k]);
  move_mass<L, true, CL>(P, mu, th, q, pib, bot, kd, m, th0, thp, q0, qarr, cl);
#pragma unroll
  for (int j = 0; j < L - 1; j++)
    if (j >= kd) wxo[j] += m;
#pragma unroll
  for (int k = 0; k < bot; k++)
    if (k > kd && Gl[k] > 0) {
      move_mass<L, true, CL>(P, mu, th, q, pib, k, kd, m * Gl[k], tk[k], thp, qk[k], qarr, cl);
#pragma unroll
      for (int j = 0; j < L - 1; j++)
        if (j >= kd && j < k) wxo[j] += m * Gl[k];
    }
#pragma unroll
  for (int k = 0; k < L; k++)
    if (k == kd && Gl[k] > 0) q[k] = fmax(0.0, q[k] + m * Gl[k] * (qarr - qk[k]) / mu[k]);
  const double Ea = column_energy<L, true>(mu, pib, th, q, P);
#pragma unroll
  for (int k = 0; k < L; k++)
    if (k == kd) th[k] += (Ec - Ea) / (mu[k] * pib[k]);
}

struct AirIO {
  double *theta, *qv, *Ts, *ice;
  double *wx, *olr, *sw_abs, *stored, *rain, *evap, *conv, *tsurf, *rainp, *qocean, *qsw, *flash, *swalb, *firn;
  double *cloudw, *cloudo;
  double* cap;                                         // (rewritten over land with c_wet)
  const double *land, *soilcap, *lc;                   // the land mask (> 0.5 land), the bucket's capacity, the cover
  double *soil, *pond, *snow;
  double* xfer;                                        // (3 L, N) the gas tracers' mole fractions, nullable
  // (5 L - 2, N) the diagnostics rows, nullable: heating and latent heating per layer (W/m^2), the crossing's gross
  // up and down per interface, the convective mass into each layer
  double* dgo;
};

template <int L, bool MIX, bool MOIST, bool CLOUD, bool FULL>
__global__ void __launch_bounds__(BS) k_air_column(int N, CP P, AP A, int mode, const double* __restrict__ t,
                                                   const double* __restrict__ insol, const double* __restrict__ pos,
                                                   const double* __restrict__ uabs, const double* __restrict__ mu_in,
                                                   double dt,
                                                   const double* __restrict__ elev, const double* __restrict__ radf,
                                                   const double* __restrict__ swf, const double* __restrict__ ch,
                                                   AirIO o) {
  const int i = XTID;
  if (i >= N) return;
  if constexpr (!FULL) {
    // (the plain column, launched only when all of these are off: their branches compile away, and the kernel keeps
    // the registers they would take)
    P.bl_mix = 0;
    A.sat_prof = A.sat_base = A.edge_vap = A.freeze_dry = A.q_shape = A.plume_root = A.plume_deep = A.bl_q = 0;
    A.cu_dth = A.cu_sub = 0;
  }
  if (elev && A.theta_rise > 0 && elev[i] > 0) P.theta_star += A.theta_rise * elev[i];
  const double ts_base = P.theta_star;
  P.bl_cap = omexp<MIX>(dt / 86400.0);                 // (the mixed layer takes in at most a day's share per step)
  const bool blm = P.bl_mix > 0;
  const int top = 0, bot = L - 1;
  double mu[L], th[L], q[L], pb[L], pib[L], Q[L], T[L], eps[L];
  double mt = 0;
#pragma unroll
  for (int k = 0; k < L; k++) {
    mu[k] = mu_in[(size_t)k * N + i];
    th[k] = o.theta[(size_t)k * N + i];
    q[k] = MOIST ? o.qv[(size_t)k * N + i] : 0.0;
    mt += mu[k];
    Q[k] = 0;
  }
  layer_exner<L>(mu, P, pb, pib);
  const double exb = pow(pb[bot] / P.p0, P.kappa);
  // the mixed bottom layer: the lowest interface's theta* no colder than the layer, below the air above's mean
  if (blm && L > 2) P.theta_star = fmin(fmax(ts_base, th[L - 1]), th[L - 2] - 1);
  const double ths_cl = L > 4 ? P.ths[L > 4 ? L - 4 : 0] : 0.0;   // (cl_float: the cloud layer's fixed lower edge)
  if constexpr (FULL) cu_edges<L, MIX>(mu, th, q, P, A, ths_cl);
  // (the diagnostics rows: each accumulates in the caller's array at this column, in the oracle's order)
  if (o.dgo)
    for (int r = 0; r < 5 * L - 2; r++) o.dgo[(size_t)r * N + i] = 0;
  auto dgl = [&](int k, double v) {
    if (o.dgo) o.dgo[(size_t)(L + k) * N + i] += v;
  };
  auto cin = [&](int k, double v) {
    if (o.dgo) o.dgo[(size_t)(4 * L - 2 + k) * N + i] += v;
  };
  // (the column's gas tracers, carried by every move of its air)
  double xfl[3 * L];
  if (o.xfer) {
    for (int g = 0; g < 3; g++)
#pragma unroll
      for (int k = 0; k < L; k++) xfl[g * L + k] = o.xfer[((size_t)g * L + k) * N + i];
    P.xf = xfl;
  }
  double m_ice = MOIST ? o.ice[i] : 0.0;
  const bool isl = MOIST && FULL && o.land && o.land[i] > 0.5;   // land: the soil bucket, ponds, snow and the ground's heat
  double* Ts = o.Ts;
  if (o.cap && A.c_wet > 0 && isl) {                     // (c_wet: the ground's capacity by its water, its heat kept)
    const double r = fmin(1.0, o.soil[i] / o.soilcap[i]), Cn = A.c_dry + (A.c_wet - A.c_dry) * r;
    Ts[i] = A.Tf + (Ts[i] - A.Tf) * o.cap[i] / Cn;
    o.cap[i] = Cn;
  }
  const double Cw = o.cap ? o.cap[i] : A.c_slab;
  const double Cf = o.firn && A.firn > 0 ? A.c_ice * A.rho_firn * A.firn : 0.0;
  double Tfn = Cf > 0 ? o.firn[i] : A.Tf;
  const double E0 = column_energy<L, MOIST>(mu, pib, th, q, P) + Cw * Ts[i] - A.Lf * m_ice + Cf * (Tfn - A.Tf);
#pragma unroll
  for (int k = 0; k < L; k++) T[k] = pib[k] * th[k] / P.cp;
  using F = std::conditional_t<MIX, float, double>;
  // clouds: each layer's cover, in-cloud water path (kg/m^2) and liquid share (all ice below -liq_cold C); a middle
  // layer's clouds sit high in it, so their phase is by its upper quarter's pressure
  double cl[L], fc[L], liq[L], cwp[L];
  if constexpr (CLOUD) {
#pragma unroll
    for (int k = 0; k < L; k++) {
      cl[k] = o.cloudw[(size_t)k * N + i];
      const double ptk = k ? pb[k - 1] : 0.0;
      const double pc = k == bot ? 0.5 * (pb[k] + ptk) : ptk + 0.25 * (pb[k] - ptk);
      liq[k] = fmax(0.0, fmin(1.0, (th[k] * fpow_<MIX>(pc / P.p0, P.kappa) - A.Tf + A.liq_cold) / A.liq_cold));
      F qs = 0, dqs;
      double l, f, dqm;
      if (!mixed_cloud<L, MIX>(k, mu, th[k], q[k] + cl[k], P, A, l, f, dqm, false)) {
        layer_cap_k<L, MIX>(k, th[k], ptk, ptk + P.grav * mu[k], P, A, qs, dqs);
        cloud_pdf(q[k] + cl[k], qs, A, l, f);
      } else if (A.cw_qs > 0) {
        layer_cap_k<L, MIX>(k, th[k], ptk, ptk + P.grav * mu[k], P, A, qs, dqs);
      }
      fc[k] = fmax(f, cover_of(cl[k], mu[k], liq[k], A, qs));
      cwp[k] = fc[k] > 1e-6 ? mu[k] * cl[k] / fc[k] : 0.0;
    }
    // cu_cloud (with ml_clear and bl_q): the mixed layer clear below its cloud base, the cloud its thermals make is
    // above it, in the layer above: the thermals' statistics at the mixed layer's levels give its cover there and its
    // water (cu_cloud 2: the cover alone), for the radiation
    if (A.cu_cloud > 0 && A.ml_clear > 0 && A.bl_q > 0 && L > 2) {
      double l_, f_, d_;
      if (mixed_cloud<L, MIX>(bot, mu, th[bot], q[bot] + cl[bot], P, A, l_, f_, d_, true) && f_ > 1e-6) {
        constexpr int kc = bot - 1;
        const double w_ = mu[kc] * cl[kc] + (A.cu_cloud > 1.5 ? 0.0 : mu[bot] * l_);
        fc[kc] = fmax(fc[kc], f_);
        cwp[kc] = w_ / fc[kc];
      }
    }
  }
  // the air touching the ground: the bottom mu_bl of the column (bl_bot: the thin bottom layer's share of it); the
  // layers above give their air at their lower edge's theta* where warmer
  double wbl[L], thb = 0, qb = 0, mb = 0;
  {
    double need = blm ? mu[bot] : A.bl_bot > 0 && mu[bot] < P.mu_bl ? fmax(mu[bot], A.bl_bot * P.mu_bl) : P.mu_bl;
#pragma unroll
    for (int k = bot; k >= 0; k--) {
      const double m = mu[k] < need ? mu[k] : need;
      wbl[k] = m;
      need -= m;
      thb += m * (A.bl_bot > 0 && k < bot && th[k] > tstar<L>(k, P) ? tstar<L>(k, P) : th[k]);
      if constexpr (MOIST) qb += m * q[k];
      mb += m;
    }
  }
  thb /= mb;
  qb /= mb;
  const double ps = pb[bot];
  const double Tsa = thb * exb;
  const bool sfc = A.sfc > 0 && ch;
  const double rough = !sfc && isl ? A.land_exch : 1.0;
  double stab = 1;
  if (A.bulk_c > 0 && A.bulk_z > 0) {
    const double Ub = uabs[i] + P.bulk_gust, Ri = P.grav * A.bulk_z * (Tsa - Ts[i]) / (Tsa * Ub * Ub);
    stab = Ri > 0 ? 1 / (1 + 10 * Ri * (1 + 5 * Ri)) : 1 - 10 * Ri / (1 + 10 * sqrt(-Ri));
  }
  double bulk = sfc ? ch[i] : A.bulk_c > 0 ? ps / (A.rd * Tsa) * A.bulk_c * stab * (uabs[i] + P.bulk_gust) : 0.0;
  if (!sfc && A.bulk_c > 0 && A.bulk_w > 0) {
    const double gt = ps / (A.rd * Tsa) * A.bulk_w * (Ts[i] > Tsa ? 3.0 : 1.0);
    bulk = bulk * gt / (bulk + gt);
  }
  if (sfc) bulk = bulk / (1 + bulk * dt / mb);
  const double hs = rough * (A.bulk_c > 0 || sfc ? P.cp * bulk : A.h_sens * (1 + uabs[i] / A.u0));

  // ---- shortwave; the sea's albedo with the ice's cover ----
  double qrad[L], qet[L], qeb[L];
#pragma unroll
  for (int k = 0; k < L; k++) qrad[k] = qet[k] = qeb[k] = 0;
  double I;
  if (mode == 0) {
    double fd;
    const double ang = sun_lon_fd(t[0], A, &fd);
    const double dec = asin(sin(A.tilt) * sin(ang));
    const double rot = 2 * 3.14159265358979323846 * fmod(t[0], A.day_s) / A.day_s;
    const double sx = cos(dec) * cos(-rot), sy = cos(dec) * sin(-rot), sz = sin(dec);
    const double cz = pos[3 * i] * sx + pos[3 * i + 1] * sy + pos[3 * i + 2] * sz;
    I = A.S0 * fd * (cz > 0 ? cz : 0.0);
  } else {
    I = insol[i];
  }
  const double dice = m_ice / P.rho_ice;
  const double fice = MOIST ? fmin(1.0, dice / P.ice_cover) : 0.0;
  // the ground's albedo by its wetness (and the cover's lc rows: 1 dry, 2 wet, 3 the vegetated share, 4 the
  // vegetation's), its ponds at the sea's; the ice's, on land by its snow; lc row 5 the share of ice that shows
  double alb = A.alb;
  const double ice0 = m_ice;
  double wet = isl ? o.soil[i] : 0.0, pw = isl ? o.pond[i] : 0.0;
  const double pcov = isl && pw > 0 ? fmin(1.0, pw / A.pond_cover) : 0.0;
  const double* lc = o.lc;
  const size_t Ns = (size_t)N;
  if constexpr (MOIST) {
    const double scap = isl ? o.soilcap[i] : 1.0;
    const double alb0 = !isl ? A.alb
                        : lc ? pcov * A.alb + (1 - pcov) * (lc[3 * Ns + i] * (lc[Ns + i] + (lc[2 * Ns + i] - lc[Ns + i]) *
                                                                                           wet / scap) +
                                                            (1 - lc[3 * Ns + i]) * lc[4 * Ns + i])
                             : pcov * A.alb + (1 - pcov) * (A.alb_dry + (A.alb_wet - A.alb_dry) * wet / scap);
    const double warm = fmax(0.0, fmin(1.0, (Tsa - A.Tf + A.alb_ramp) / A.alb_ramp));
    double alb_i = A.alb_ice + (A.alb_ice_melt - A.alb_ice) * warm;
    if (isl) {
      const double fs = fmin(1.0, o.snow[i] / A.snow_cover);
      alb_i = fs * (A.alb_snow + (A.alb_snow_melt - A.alb_snow) * warm) + (1 - fs) * A.alb_glacier;
    }
    const double fsee = isl && lc ? fice * lc[5 * Ns + i] : fice;
    alb = fsee * alb_i + (1 - fsee) * alb0;
  }
  if (o.swalb) o.swalb[i] = alb;
  double toAir;
  if (swf) {
    double s = 0;
#pragma unroll
    for (int k = 0; k < L; k++) s += swf[(size_t)k * N + i];
    toAir = s * I;
#pragma unroll
    for (int k = 0; k < L; k++) {
      const double f = swf[(size_t)k * N + i];
      Q[k] += I * f;
      qrad[k] += I * f;
      qet[k] += I * swf[(size_t)(L + 2 + 2 * k) * N + i];
      qeb[k] += I * swf[(size_t)(L + 3 + 2 * k) * N + i];
    }
  } else {
    toAir = A.asw * (mt / A.mu_ref) * I;
    if (A.sw_tau > 1) {
      double st = 0;
#pragma unroll
      for (int k = 0; k < L; k++) {
        const double sb = pb[k] / pb[bot], sb4 = sb * sb * sb * sb;
        Q[k] += toAir * (sb4 - st);
        st = sb4;
      }
    } else if (A.sw_tau > 0) {
#pragma unroll
      for (int k = 0; k < L; k++)
        Q[k] += toAir * (tau_p(pb[k], pb[bot], A) - tau_p(k ? pb[k - 1] : 0.0, pb[bot], A)) / A.tau0;
    } else {
#pragma unroll
      for (int k = 0; k < L; k++) Q[k] += toAir * mu[k] / mt;
    }
  }
  // clouds reflect: each layer's by conservative two-stream scattering of its optical depth 3 W / (2 rho r), the
  // layers at random overlap, with the reflections between them and the ground
  double Tc = 1;
  if constexpr (CLOUD) {
#pragma unroll
    for (int k = 0; k < L; k++) {
      const double tau = 1.5 * cwp[k] * (liq[k] / (1000 * A.re_liq) + (1 - liq[k]) / (P.rho_ice * A.re_ice));
      const double g1 = (1 - A.cloud_g) * tau;
      Tc *= 1 - fc[k] * g1 / (2 + g1);
    }
  }
  const double swS = swf ? I * swf[(size_t)L * N + i] : (I - toAir) * Tc * (1 - alb) / (1 - alb * (1 - Tc));

  // ---- longwave ----
  if (MOIST && A.lw_w0 > 0) {
    double W = 0;
#pragma unroll
    for (int k = 0; k < L; k++) W += mu[k] * (q[k] > 0 ? q[k] : 0.0);
    const double tv = A.tau0 * (1 - A.tau_lin) * sqrt(W / A.lw_w0);
#pragma unroll
    for (int k = 0; k < L; k++) {
      const double sb = pb[k] / pb[bot], st = k ? pb[k - 1] / pb[bot] : 0.0;
      eps[k] = omexp<MIX>((A.tau0 * A.tau_lin * (sb - st) + (W > 0 ? tv * mu[k] * (q[k] > 0 ? q[k] : 0.0) / W : 0.0)));
    }
  } else {
#pragma unroll
    for (int k = 0; k < L; k++)
      eps[k] = omexp<MIX>(tau_p(pb[k], pb[bot], A) - tau_p(k ? pb[k - 1] : 0.0, pb[bot], A));
  }
  double eps0[L];
  if constexpr (CLOUD) {                             // clouds absorb and emit by their water path
#pragma unroll
    for (int k = 0; k < L; k++) {
      eps0[k] = eps[k];
      const double kl = liq[k] * A.k_lw_liq + (1 - liq[k]) * A.k_lw_ice;
      eps[k] = 1 - (1 - eps[k]) * (1 - fc[k] * omexp<MIX>(kl * cwp[k]));
    }
  }
  double ldown = 0;
  constexpr int RW = L + 1;
  if (radf) {
    ldown = radf[(size_t)(2 * RW - 1) * N + i] - radf[(size_t)L * N + i];
  } else {
#pragma unroll
    for (int k = 0; k < L; k++) {
      double dn = eps[k] * A.sigma * T[k] * T[k] * T[k] * T[k];
#pragma unroll
      for (int j = k + 1; j < L; j++) dn -= eps[j] * dn;
      ldown += dn;
    }
  }
  // ---- the surface: open water at Ts, or under ice thicker than ice_cover a skin balancing conduction through the
  // ice against the surface's fluxes, capped at freezing ----
  const double kex = rough * (A.bulk_c > 0 || sfc ? bulk : A.k_evap * (1 + uabs[i] / A.u0));
  // q_shape: the air at the surface holds the bottom layer's vapor at its relative humidity there (against the
  // layer's capacity), not the layer's mean; bl_q keeps the layer's vapor even
  double qa = qb;
  if (MOIST && A.q_shape > 0 && !(A.bl_q > 0)) {
    F qsl, dqsl;
    layer_cap_k<L, MIX>(bot, th[bot], pb[bot - 1], pb[bot], P, A, qsl, dqsl);
    const double qss = (double)qsat_a<F>((F)Tsa, (F)ps, (F)A.eps, 0.0);
    if (qsl > 0) qa = fmin(qb * qss / qsl, qss);
  }
  // soil and plants hold their water back: a surface resistance rs in series with the air's ra (land_rs; rs_max: the
  // stomata open with the sunlight, rs = land_rs (1 + f) / (land_rs / rs_max + f), f the sunlight reaching the ground
  // over rs_sun; lc rows 6, 7 the cover's own land_rs, rs_sun; its bare share, row 3's complement, has none);
  // bare_fx: the bare soil's evaporation falls as (wet / full)^bare_fx, the plants' as wet / full
  double rfac = 1, fbare = 0, rveg = 1;
  if (isl && A.land_rs > 0 && kex > 0) {
    const double r0 = lc ? lc[6 * Ns + i] : A.land_rs, rsun = lc ? lc[7 * Ns + i] : A.rs_sun;
    double rs = r0;
    if (A.rs_max > 0) {
      const double f = swS / fmax(1 - alb, 0.05) / rsun;
      rs = r0 * (1 + f) / (r0 / A.rs_max + f);
    }
    const double ra = ps / (A.rd * Tsa) / kex;
    rfac = ra / (ra + rs);
    rveg = rfac;
    if (lc) {
      fbare = lc[3 * Ns + i];
      rfac = fbare + (1 - fbare) * rfac;
    }
  }
  const bool bfx = isl && A.bare_fx > 0 && lc && A.land_rs > 0;
  auto soilf = [&](double w_) {
    const double x_ = fmin(1.0, w_);
    return fbare * pow(x_, A.bare_fx) + (1 - fbare) * rveg * x_;
  };
  const double qfac = !isl && P.q_sea > 0 ? P.q_sea : 1.0;
  const bool skin = MOIST && dice > P.ice_cover;
  // firn where the land's ice is thick enough; where it ends, its heat goes to the ground (or the slab) below
  const bool fon = Cf > 0 && isl && skin && m_ice >= A.rho_firn * A.firn;
  if (Cf > 0 && !fon && Tfn != A.Tf) {
    Ts[i] += Cf * (Tfn - A.Tf) / Cw;
    Tfn = A.Tf;
  }
  const double ts = Ts[i];
  double tsk = ts, melt_top = 0, cond = 0, condb = 0;
  if (skin) {
    // the skin conducts to what lies under it: the ice down to the ground (or the slab), or the firn's middle
    const double kc = fon ? 2 * A.k_firn / A.firn : A.k_ice / dice, tb = fon ? Tfn : ts;
    double x = fmin(tb, A.Tf);
    for (int it = 0; it < 8; it++) {
      const double x3 = x * x * x;
      const double f = kc * (tb - x) - A.eps_s * A.sigma * x3 * x + swS + ldown - hs * (x - Tsa);
      x -= f / (-kc - 4 * A.eps_s * A.sigma * x3 - hs);
    }
    if (x > A.Tf) x = A.Tf;
    const double x4 = x * x * x * x;
    melt_top = kc * (tb - x) - A.eps_s * A.sigma * x4 + swS + ldown - hs * (x - Tsa);
    tsk = x;
    cond = kc * (tb - tsk);
    if (fon) {                                       // and from the ground through the rest of the ice to the firn
      const double rest = fmax(dice - A.rho_firn * A.firn / P.rho_ice, 0.0);
      condb = (ts - Tfn) / (rest / A.k_ice + A.firn / (2 * A.k_firn));
    }
  }
  // the ground's skin (c_deep; its capacity Cw): stepped implicitly, C (x - ts) = dt F(x) with F the surface's net
  // gain (sunlight, longwave, sensible heat, the evaporation before the water runs short), by Newton's method
  if (isl && !skin && A.c_deep > 0) {
    const double wfull = wet / fmax(A.soil_beta * o.soilcap[i], 1e-9);
    const double beta = kex * (pcov + (1 - pcov) * (bfx ? soilf(wfull) : rfac * fmin(1.0, wfull))) * (1 - fice);
    double x = ts;
    for (int it = 0; it < 6; it++) {
      const double x3 = x * x * x, d = (double)qsat_a<F>((F)x, (F)ps, (F)A.eps, 0.0) - qa;
      const double dqs = (double)(qsat_a<F>((F)(x + 0.01), (F)ps, (F)A.eps, 0.0) -
                                  qsat_a<F>((F)(x - 0.01), (F)ps, (F)A.eps, 0.0));
      const double e = d > 0 ? beta * d : 0.0, de = d > 0 ? beta * dqs * 50 : 0.0;
      const double g = Cw * (x - ts) - dt * (swS + ldown - A.eps_s * A.sigma * x3 * x - hs * (x - Tsa) - P.Lv * e);
      x -= g / (Cw + dt * (4 * A.eps_s * A.sigma * x3 + hs + P.Lv * de));
    }
    tsk = x;
  }
  double up = A.eps_s * A.sigma * tsk * tsk * tsk * tsk;
  if (radf) {
    const double Es = up;
    double n0 = radf[i] - radf[(size_t)RW * N + i], tr0 = radf[(size_t)(2 * RW) * N + i];   // edge 0's net, transmission
    const double up0 = radf[i];
#pragma unroll
    for (int k = 0; k < L; k++) {
      const double n1e = radf[(size_t)(k + 1) * N + i] - radf[(size_t)(RW + k + 1) * N + i];
      const double tr1 = radf[(size_t)(2 * RW + k + 1) * N + i];
      const double ql = n1e - n0 + (tr1 - tr0) * Es;
      Q[k] += ql;
      const double* re = radf + (size_t)(3 * RW + 3 + 6 * k) * N + i;
      const double s1 = re[0] - re[N] + re[2 * (size_t)N] * Es;
      const double s2 = re[3 * (size_t)N] - re[4 * (size_t)N] + re[5 * (size_t)N] * Es;
      qrad[k] += ql;
      qet[k] += s1 - (n0 + tr0 * Es);
      qeb[k] += (n1e + tr1 * Es) - s2;
      n0 = n1e;
      tr0 = tr1;
    }
    up = up0 + radf[(size_t)(2 * RW) * N + i] * Es;
    if constexpr (CLOUD) {
      if (o.cloudo) {                                // cover, the clouds' effects at the top, optical depths
        double* co = o.cloudo + i;
        double clear = 1;
#pragma unroll
        for (int k = 0; k < L; k++) {
          co[(size_t)k * N] = fc[k];
          clear *= 1 - fc[k];
        }
        co[(size_t)L * N] = 1 - clear;
        co[(size_t)(L + 1) * N] =
            swf ? I * (swf[(size_t)(L + 1) * N + i] - (1 - (toAir / fmax(I, 1e-30) + swf[(size_t)L * N + i])))
                : swS - (I - toAir) * (1 - alb);
        co[(size_t)(L + 2) * N] = radf[(size_t)(3 * RW) * N + i] + radf[(size_t)(3 * RW + 1) * N + i] * Es - up;
        const double re_l = 1000 * A.re_liq, re_i = P.rho_ice * A.re_ice;
#pragma unroll
        for (int k = 0; k < L; k++) {
          co[(size_t)(L + 3 + k) * N] = fc[k] > 1e-6 ? (liq[k] * A.k_lw_liq + (1 - liq[k]) * A.k_lw_ice) * cwp[k] : 0.0;
          co[(size_t)(2 * L + 3 + k) * N] = 1.5 * cwp[k] * (liq[k] / re_l + (1 - liq[k]) / re_i);
          const double tl_ = liq[k] / re_l, ti_ = (1 - liq[k]) / re_i;
          co[(size_t)(3 * L + 3 + k) * N] = tl_ + ti_ > 0 ? tl_ / (tl_ + ti_) : 1.0;
        }
      }
    }
  } else {
#pragma unroll
    for (int k = bot; k >= 0; k--) {
      const double a = eps[k] * up;
      Q[k] += a;
      const double em = eps[k] * A.sigma * T[k] * T[k] * T[k] * T[k];
      Q[k] -= 2 * em;
      up = up - a + em;
    }
    if constexpr (CLOUD) {
      if (o.cloudo) {                                // cover and the clouds' effects at the top
        double* co = o.cloudo + i;
        double up0 = A.eps_s * A.sigma * tsk * tsk * tsk * tsk, clear = 1;
#pragma unroll
        for (int k = bot; k >= 0; k--)
          up0 = up0 - eps0[k] * up0 + eps0[k] * A.sigma * T[k] * T[k] * T[k] * T[k];
#pragma unroll
        for (int k = 0; k < L; k++) {
          co[(size_t)k * N] = fc[k];
          clear *= 1 - fc[k];
        }
        co[(size_t)L * N] = 1 - clear;
        co[(size_t)(L + 1) * N] = swS - (I - toAir) * (1 - alb);
        co[(size_t)(L + 2) * N] = up0 - up;
      }
    }
#pragma unroll
    for (int k = 0; k < L; k++) {
      double dn = eps[k] * A.sigma * T[k] * T[k] * T[k] * T[k];
#pragma unroll
      for (int j = k + 1; j < L; j++) {
        const double a = eps[j] * dn;
        Q[j] += a;
        dn -= a;
      }
    }
  }
  o.olr[i] = up;
  // ---- sensible heat and evaporation into the bottom air; the heating as theta ----
  const double H = hs * (tsk - Tsa);
#pragma unroll
  for (int k = 0; k < L; k++) Q[k] += H * wbl[k] / mb;
  double E = 0;
  if constexpr (MOIST) {
    using F = std::conditional_t<MIX, float, double>;
    const double def = qfac * qsat_a<F>((F)tsk, (F)ps, (F)A.eps, 0.0) - qa;
    E = def > 0 ? kex * def * (1 - fice) : 0.0;
    if (isl) {                                       // land: from its ponds freely, from the soil held back
      const double full = A.soil_beta * o.soilcap[i];
      double Es = (1 - pcov) * E, Ep = pcov * E;
      if (bfx) {
        Es *= soilf(wet / full);
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
#pragma unroll
    for (int k = 0; k < L; k++) q[k] += E * dt * wbl[k] / mb / mu[k];
  }
  o.evap[i] = E;
  o.tsurf[i] = tsk;
  double dth[L], dlat[L];
#pragma unroll
  for (int k = 0; k < L; k++) {
    dth[k] = Q[k] * dt / (mu[k] * pib[k]);
    th[k] += dth[k];
    dlat[k] = 0;
  }
  // ---- the slab: open water takes the surface's fluxes (evaporation costs it Lv per kg), under ice the conduction
  // up into the skin; below Tf its heat freezes into ice, ice melts with slab heat above Tf and from the top ----
  if (fon) Tfn += (condb - cond) * dt / Cf;         // the firn between the ground's conduction and the skin's
  const double qS = skin ? (fon ? -condb : -cond) : swS + ldown - A.eps_s * A.sigma * tsk * tsk * tsk * tsk - H - P.Lv * E;
  double Tn = ts + qS * dt / Cw;
  if (o.qocean) o.qocean[i] = qS;
  if (o.qsw) o.qsw[i] = skin ? 0.0 : swS;
  if constexpr (MOIST) {
    m_ice -= melt_top * dt / A.Lf;
    if (m_ice < 0) {
      Tn += -m_ice * A.Lf / Cw;
      m_ice = 0;
    }
    if (Tn < A.Tf) {
      if (!isl) {                                    // (the ground cools below freezing; the sea freezes)
        m_ice += (A.Tf - Tn) * Cw / A.Lf;
        Tn = A.Tf;
      }
    } else if (m_ice > 0) {
      const double m = (Tn - A.Tf) * Cw / A.Lf;
      if (m <= m_ice) {
        m_ice -= m;
        Tn = A.Tf;
      } else {
        Tn = A.Tf + (m - m_ice) * A.Lf / Cw;
        m_ice = 0;
      }
    }
    o.ice[i] = m_ice;
  }
  Ts[i] = Tn;
  o.sw_abs[i] = toAir + swS;
  // ---- condensation: one-step saturation adjustment per layer, its latent heat heating for the crossing ----
  double R = 0, Rk0 = 0, Rk1 = 0, Rc = 0, S = 0;
  if constexpr (CLOUD) {
    // cloud water forms and erodes; liquid turns to rain (Sundqvist 1989), ice falls out; rain evaporates in the clear
    // air it falls through. With ice_air the falling ice freezes in its layer (Lf), melts in a layer whose lower edge
    // is above freezing, sublimates with the rain's evaporation; what is left reaches the ground as snow.
    const bool icy = A.ice_air > 0;
    double Fr = 0, Fs = 0, fa = 0, psf = 0;
#pragma unroll
    for (int k = 0; k < L; k++) psf += P.grav * mu[k];
#pragma unroll
    for (int k = 0; k < L; k++) {
      double pt = 0;
#pragma unroll
      for (int j = 0; j < k; j++) pt += P.grav * mu[j];
      const double pbk = pt + P.grav * mu[k];
      // condense up to the statistics' condensate with the latent heat's feedback on the capacity; a cloud holding
      // more erodes over tau_erode (air.cu:438-450)
      F qs, dqs;
      double l, f, dl = 0, dqa = 0, dqm;
      bool go;
      if (mixed_cloud<L, MIX>(k, mu, th[k], q[k] + cl[k], P, A, l, f, dqm, false)) {
        f = 1;                                       // (dqm already carries each level's share)
        dqa = dqm;
        go = true;
      } else {
        layer_cap_k<L, MIX>(k, th[k], pt, pbk, P, A, qs, dqs);
        go = qs > 0;
        if (go) {
          cloud_pdf(q[k] + cl[k], qs, A, l, f);
          dqa = dqs;
        }
      }
      if (go) {
        dl = l > cl[k] ? (l - cl[k]) / (1 + P.Lv / pib[k] * dqa * f) : -(cl[k] - l) * omexp<MIX>(dt / A.tau_erode);
        cl[k] += dl;
        q[k] -= dl;
        th[k] += P.Lv * dl / pib[k];
      }
      dth[k] += P.Lv * dl / pib[k];
      dlat[k] += P.Lv * dl / pib[k];
      dgl(k, P.Lv * mu[k] * dl / dt);
      layer_cap_k<L, MIX>(k, th[k], pt, pbk, P, A, qs, dqs);
      if (!mixed_cloud<L, MIX>(k, mu, th[k], q[k] + cl[k], P, A, l, f, dqm, false)) cloud_pdf(q[k] + cl[k], qs, A, l, f);
      f = fmax(f, cover_of(cl[k], mu[k], liq[k], A, qs));
      const double inc = f > 1e-6 ? cl[k] / f * fmax(1.0, mu[k] / A.mu_cloud) : 0.0, r = inc / A.q_auto;
      const double pm = pt + 0.5 * P.grav * mu[k], Tk = pib[k] * th[k] / P.cp, depth = mu[k] / (pm / (A.rd * Tk));
      const double rate = liq[k] * A.c_auto * omexp<MIX>(r * r) + (1 - liq[k]) * A.v_ice / depth;
      const double loss = cl[k] * omexp<MIX>(rate * dt);
      cl[k] -= loss;
      Fr += mu[k] * loss / dt;
      if (icy) {
        const double fz = rate > 0 ? fmin(1.0, (1 - liq[k]) * A.v_ice / depth / rate) : 0.0, dz = loss * fz;
        Fs += mu[k] * dz / dt;
        th[k] += A.Lf * dz / pib[k];
        dth[k] += A.Lf * dz / pib[k];
        dlat[k] += A.Lf * dz / pib[k];
        dgl(k, A.Lf * mu[k] * dz / dt);
        if (Fs > 0 && base_T<L, MIX>(k, th[k], pbk, P, A) > A.Tf) {          // (the layer's lower edge above freezing)
          const double dm = Fs * dt / (mu[k] * pib[k]) * A.Lf;
          th[k] -= dm;
          dth[k] -= dm;
          dlat[k] -= dm;
          dgl(k, -(A.Lf * Fs));
          Fs = 0;
        }
      }
      double ev = 0;
      if (loss > 0) fa = fmax(fa, f);
      const double rc_ = A.revap > 0 ? (isl ? A.revap_rhl : A.revap_rhs) : 1.0;
      if (Fr > 0 && rc_ * qs > q[k]) {
        if (A.revap > 0) {
          const double a_ = fmax(fa, A.revap_area), ac = fmax(0.0, a_ - f);
          ev = ac * 5.44e-4 * (rc_ * qs - q[k]) * pow(sqrt(pm / psf) / 5.09e-3 * Fr / a_, 0.577) * mu[k];
        } else {
          ev = A.k_revap * (1 - f) * (1 - q[k] / qs) * sqrt(Fr) * mu[k];
        }
        ev = fmin(ev, fmin(Fr, (rc_ * qs - q[k]) * mu[k] / dt / (1 + P.Lv / pib[k] * dqs)));
        const double es = Fr > 0 ? ev * Fs / Fr : 0.0;
        const double Le = P.Lv * ev + A.Lf * es;
        Fr -= ev;
        Fs -= es;
        q[k] += ev * dt / mu[k];
        th[k] -= Le * dt / (mu[k] * pib[k]);
        dth[k] -= Le * dt / (mu[k] * pib[k]);
        dlat[k] -= Le * dt / (mu[k] * pib[k]);
        dgl(k, -Le);
      }
      if (k == bot) Rk1 += mu[k] * loss - ev * dt;
      else Rk0 += mu[k] * loss - ev * dt;
    }
    R = Fr * dt;
    S = fmax(0.0, fmin(Fs * dt, R));
  } else if constexpr (MOIST) {
#pragma unroll
    for (int k = 0; k < L; k++) {
      double tk = th[k], qk = q[k];
      double pt = 0;
#pragma unroll
      for (int j = 0; j < k; j++) pt += P.grav * mu[j];
      const double dq = condense_at<L, MIX>(k, tk, qk, pib[k], pt, mu[k], P, A);
      th[k] = tk;
      q[k] = qk;
      dth[k] += P.Lv * dq / pib[k];
      dlat[k] += P.Lv * dq / pib[k];
      dgl(k, P.Lv * mu[k] * dq / dt);
      R += mu[k] * dq;
      if (k == bot) Rk1 += mu[k] * dq;
      else Rk0 += mu[k] * dq;
    }
  }
  // ---- the crossing: the k-tables' edge sub-levels heat the layers' edges ----
  double et[L], eb[L], wxo[L];
#pragma unroll
  for (int k = 0; k < L; k++) {
    et[k] = dth[k];
    eb[k] = dth[k];
  }
  eb[0] = dth[top];
  if (radf) {
    const bool edges = A.edge_heat > 0;
    const bool de = A.dry_edge > 0 && L > 2;
    double pt_ = 0;
#pragma unroll
    for (int k = 0; k < L; k++) {
      const double pbk = pt_ + P.grav * mu[k], dp = pbk - pt_;
      const double ft = k == 0 ? 0.05 : 0.25, fb = k == 0 ? 0.5 : 0.25;
      const double pit = P.cp * fpow_<MIX>((pt_ + 0.5 * ft * dp) / P.p0, P.kappa);
      const double pibt = P.cp * fpow_<MIX>((pbk - 0.5 * fb * dp) / P.p0, P.kappa);
      const double mean = qrad[k] * dt / (mu[k] * pib[k]);
      double dtop = dth[k] - mean + qet[k] * dt / (ft * mu[k] * pit);
      double dbot = dth[k] - mean + qeb[k] * dt / (fb * mu[k] * pibt);
      if (k == L - 1 && blm) dtop = dbot = dth[k];    // (the mixed bottom layer heats evenly)
      if (de && k == (int)A.edge_k) dtop -= dlat[k];
      if (edges) {
        et[k] = dtop;
        eb[k] = dbot;
      } else if (k >= 1 && k <= A.edge_k) {
        et[k] = dtop;
      }
      pt_ = pbk;
    }
  }
  if (blm && L > 2) P.theta_star = fmin(fmax(ts_base, th[bot]), th[bot - 1] - 1);
  if constexpr (FULL) cu_edges<L, MIX>(mu, th, q, P, A, ths_cl);
  double mu_lcl = 1e30;
  if (A.bl_lcl > 0 && blm && L > 2) {                  // (bl_lcl: the mixed layer ends at its condensation level)
    mu_lcl = lcl_mass<L, MIX>(mu, th[bot], q[bot], P, A);
    P.bl_room = fmax(0.0, mu_lcl - mu[bot]);
  }
  crossing<L, MOIST, CLOUD, MIX>(P, mu, th, q, dth, et, eb, pib, wxo, cl, &A,
                                 o.dgo ? o.dgo + (size_t)2 * L * N + i : nullptr, (size_t)N);
  // the mixed boundary layer: the surface's buoyancy flux sets its depth. Rising it entrains the air above at bl_ent
  // of that flux across the jump to it (Tennekes 1973), a slice from the layer's bottom edge; falling, it gives its
  // top back down to a thin night layer mu_night. bl_reach: dry convection first takes in the air above colder than
  // it, layer by layer while they empty. bl_lcl: it ends at its condensation level, the air above that leaving into
  // the layer above.
  if (blm && L > 2) {
    auto reach_edge = [&](int k) { return k < bot - 1 ? fmax(tstar<L>(k, P), th[bot]) : tstar<L>(k, P); };
    int ka = bot - 1;
    if (A.bl_reach > 0) {
      bool stop = false;
#pragma unroll
      for (int k = bot - 1; k >= 1; k--) {
        if (stop) continue;
        const double avail = mu[k] - P.floor_;
        if (avail > 0) {
          const double e = reach_edge(k), D = th[k] - e;
          const bool spread = D > 0 && th[k] < tstar<L>(k - 1, P);
          const double x = spread ? (th[bot] > e ? mu[bot] / mu[k] * (sqrt(1 + mu[k] * (th[bot] - e) / (mu[bot] * D)) - 1)
                                                 : 0.0)
                                  : (th[k] < th[bot] ? 1.0 : 0.0);
          const double want = fmin(1.0, x) * mu[k];
          const double m = fmin(fmin(fmin(want, avail), P.bl_cap * fmax(mu[k] - P.mu_bl, P.mu_bl)),
                                fmax(0.0, mu_lcl - mu[bot]));
          if (m > 0) {
            const double fm = m / mu[k], thl = spread ? e + fm * D : th[k];
            const double qx = A.edge_vap > 0 ? edge_q<L, MIX>(k, false, mu, th, q, m, P, A) : q[k];
            move_mass<L, MOIST, CLOUD>(P, mu, th, q, pib, k, bot, m, thl, thl, qx, qx, cl);
#pragma unroll
            for (int j = k; j < bot; j++) wxo[j] -= m;
          }
          if (!(m >= avail * (1 - 1e-9))) stop = true;
        }
      }
      ka = bot - 1;
#pragma unroll
      for (int k = bot - 1; k >= 2; k--)                // (the lowest layer above with air, not below layer 1)
        if (ka == k && !(mu[k] > P.floor_ * 1.000001)) ka = k - 1;
    }
    double mka = 0, thka = 0, qka = 0, eka = 0, tska = 0;
#pragma unroll
    for (int k = 0; k < L; k++)
      if (k == ka) {
        mka = mu[k];
        thka = th[k];
        if constexpr (MOIST) qka = q[k];
        eka = reach_edge(k);
        tska = k >= 1 ? tstar<L>(k - 1, P) : 0.0;
      }
    const double Bf = H + A.vf * P.cp * Tsa * E, jump = fmax(1.0, eka - th[bot]);
    if (Bf > 0) {
      const double m = fmin(fmin(A.bl_ent * Bf * dt / (P.cp * jump), mka - P.floor_),
                            A.ent_lcl > 0 ? 1e30 : fmax(0.0, mu_lcl - mu[bot]));
      if (m > 0) {
        const bool spread = ka >= 1 && thka < tska && thka > eka;
        const double fm = fmin(1.0, m / mka);
        const double thc = spread ? eka + fm * (thka - eka) : thka;
        const double qx = MOIST && A.edge_vap > 0 ? edge_q<L, MIX>(ka, false, mu, th, q, m, P, A) : qka;
        move_mass<L, MOIST, CLOUD>(P, mu, th, q, pib, ka, bot, m, thc, thc, qx, qx, cl);
#pragma unroll
        for (int j = 0; j < bot; j++)
          if (j >= ka) wxo[j] -= m;
      }
    }
    if (!(Bf > 0) && mu[bot] > A.mu_night) {
      const double m = fmin(-Bf * dt / (P.cp * jump), mu[bot] - A.mu_night);
      const double qx = MOIST && A.edge_vap > 0 ? edge_q<L, MIX>(bot, true, mu, th, q, m, P, A) : q[bot];
      move_mass<L, MOIST, CLOUD>(P, mu, th, q, pib, bot, bot - 1, m, th[bot], th[bot], qx, qx, cl);
      wxo[bot - 1] += m;
    }
    if (A.bl_lcl > 0) {
      const double m = mu[bot] - fmax(lcl_mass<L, MIX>(mu, th[bot], q[bot], P, A), A.mu_night);
      if (m > 0) {
        const double qx = MOIST && A.edge_vap > 0 ? edge_q<L, MIX>(bot, true, mu, th, q, m, P, A) : q[bot];
        move_mass<L, MOIST, CLOUD>(P, mu, th, q, pib, bot, bot - 1, m, th[bot], th[bot], qx, qx, cl);
        wxo[bot - 1] += m;
      }
    }
  }
  // ---- moist convection: a parcel of bottom air lifted to the interface above it; buoyant there, it rises through
  // every interface it stays buoyant at and detrains above the last, its condensate raining; the mass that would
  // remove the buoyancy at the first interface relaxes over tau_conv ----
  double mc = 0, Scg = 0, qss_r = 0, ex_r = -1, ws_r = 0, rho_r = 0;   // (the root's saturation, the thermals'
                                                                         // excess, w* and density: the shallow plume's)
  const double mu_pre = mu[bot];                       // (cu_sub: the mixed layer's mass before the plumes)
  if (o.flash) o.flash[i] = 0;
  if (MOIST && P.theta_star > 0) {
    double pi_ = 0;
#pragma unroll
    for (int k = 0; k < bot; k++) pi_ += P.grav * mu[k];
    const double th0 = th[bot];
    double q0 = q[bot];
    if (A.plume_root > 0) {
      // plume_root: plumes root in the air at the surface, the layer's vapor at its relative humidity there (edge_q
      // at its bottom edge); root_q: plus the thermals' excess b E / (rho w_m), b = 8.5, w_m = (u*^3 + 0.28 w*^3)^(1/3)
      // (Holtslag & Boville 1993), up to the surface air's saturation; else with clouds its moistest part by the
      // statistics, against the surface air's capacity
      const double ps0 = pi_ + P.grav * mu[bot], Ts0 = th0 * fpow_<MIX>(ps0 / P.p0, P.kappa);
      const double qss = Ts0 < 160 ? 0.0 : (double)qsat_a<F>((F)Ts0, (F)ps0, (F)A.eps, A.ice_air);
      q0 = edge_q<L, MIX>(bot, false, mu, th, q, 0, P, A);
      if (A.root_q > 0) {
        const double rho = ps / (A.rd * Tsa), Ub = uabs[i] + P.bulk_gust;
        const double us3 = sfc ? fpow_<MIX>(fmax(0.0, ch[i] / rho * Ub), 1.5) : 0.0;
        const double Bk = fmax(0.0, (H + A.vf * P.cp * Tsa * E) / (rho * P.cp));
        const double hb = A.rd * Tsa / P.grav * log(ps / fmax(ps - P.grav * mu[bot], 1.0));
        const double wm = cbrt(us3 + 0.28 * P.grav / Tsa * Bk * hb);
        const double ex = wm > 1e-3 ? 8.5 * fmax(0.0, E) / (rho * wm) : 0.0;
        q0 += fmin(ex, fmax(0.0, qss - q0));
        qss_r = qss;
        ex_r = ex;
        rho_r = rho;
        ws_r = cbrt(P.grav / Tsa * Bk * hb);
      } else if (CLOUD) {
        q0 = fmax(q0, fmin(qss, q0 + (1 - A.rh_crit) * qss));
      }
    } else if constexpr (CLOUD) {                    // plumes root in the cell's moist air: its moistest part
      F qs, dqs;
      layer_cap_k<L, MIX>(bot, th0, pi_, pi_ + P.grav * mu[bot], P, A, qs, dqs);
      q0 = fmax(q0, fmin((double)qs, q0 + (1 - A.rh_crit) * qs));
    }
    double thp, qc, thp2, qc2;
    if (A.plume_deep > 0 && bot >= 2) {
      // the deep plume: its full ascent against the layers' spread air, mixing and entraining; the air it rises
      // through sinks to make room, warming each level by the theta it had below, so the mass that removes its
      // buoyancy is A / (theta_d - theta*), relaxed over tau_conv; it takes in the cloud layer's air it entrained (me),
      // keeps l_det of its condensate as cloud, rains the rest (snow by its frozen share), and its rain feeds a
      // downdraft of the cloud layer's air, saturated by evaporating it
      double Ab, G = 1, FZ = 0, qx[L], lt[4], xs = 0, Gl[L];
      int kd;
      plume_ascent<L, MIX>(mu, th, q, th0, q0, pi_, P, A, Ab, kd, thp, qc, qx, lt, G, FZ, Gl);
#pragma unroll
      for (int k = 0; k < L; k++) xs += qx[k];
      if (kd >= 0 && Ab > 0) {
        const double rise = thp - tstar<L>(bot - 1, P);
        mc = fmin(Ab / (rise > 1 ? rise : 1.0), mu[bot] - P.floor_) * omexp<MIX>(dt / A.tau_conv);
        if (A.plume_root > 0 && q0 > 0) mc = fmin(mc, mu[bot] * q[bot] / q0);
        // the mixed boundary layer (bl_q): plumes leave it through the clouds at its top, at most rho a w* with a
        // their share there and w* the free convection's velocity scale (Neggers, Stevens & Neelin 2006)
        if (blm && A.bl_q > 0) {
          double l_, f_, d_;
          if (mixed_cloud<L, MIX>(bot, mu, th[bot], q[bot] + (CLOUD ? cl[bot] : 0.0), P, A, l_, f_, d_,
                                  A.cu_cloud > 0)) {
            const double rho = ps / (A.rd * Tsa), Bk = fmax(0.0, (H + A.vf * P.cp * Tsa * E) / (rho * P.cp));
            const double hb = A.rd * Tsa / P.grav * log(ps / fmax(ps - P.grav * mu[bot], 1.0));
            const double ws = cbrt(P.grav / Tsa * Bk * hb);
            mc = fmin(mc, rho * f_ * ws * dt);
          }
        }
        // cu_dth: through the cumulus and cloud layers, the air it took in from each (the layer it stops in included)
        // moved by plume_out
        double cond_cu = 0, det_cu = 0, G_cu = 1;
        const bool cu = A.cu_dth > 0;
        if (mc > 0 && cu)
          plume_out<L, CLOUD>(P, mu, th, q, pib, cl, kd, mc, th0, q0, thp, qc, Gl, qx, wxo, A.l_det, cond_cu, det_cu,
                              G_cu);
        if (mc > 0) {
          if (!(kd < bot - 1)) G = 1;
          const double me = cu ? mc * (G_cu - 1) : fmin(mc * (G - 1), fmax(0.0, mu[bot - 1] - P.floor_));
          const double qe = q[bot - 1], the = th[bot - 1];
          G = cu ? G_cu : 1 + me / mc;
          const double cond_ = cu ? cond_cu : q0 + (G - 1) * qe - G * qc - xs;
          const double cond = cu ? cond_cu : fmax(0.0, cond_);
          const double det = cu ? det_cu : CLOUD ? fmin(cond, A.l_det * G) : 0.0;
          if (o.flash) o.flash[i] = storm_flashes<MIX>(mc / dt, lt, P, A);
          R += mc * (cond - det);
          Rc = mc * (cond - det);
          if (!cu) {
            const double qarr = qc + (cond_ - cond) / G;
            const double Ec = column_energy<L, MOIST>(mu, pib, th, q, P);
#pragma unroll
            for (int k = 0; k < bot; k++) q[k] = fmax(0.0, q[k] + mc * qx[k] / mu[k]);
            move_mass<L, true, CLOUD>(P, mu, th, q, pib, bot, kd, mc, th0, thp, q0, qarr, cl);
            if (me > 0) move_mass<L, true, CLOUD>(P, mu, th, q, pib, bot - 1, kd, me, the, thp, qe, qarr, cl);
            const double Ea = column_energy<L, MOIST>(mu, pib, th, q, P);
#pragma unroll
            for (int k = 0; k < L; k++)
              if (k == kd) th[k] += (Ec - Ea) / (mu[k] * pib[k]);
          }
          double Sc = 0;
          if (A.ice_air > 0 && cond > 0 && cond > det) {   // (its frozen share snows, melting below freezing air)
            Sc = mc * (cond - det) * fmin(1.0, FZ / cond);
#pragma unroll
            for (int k = 0; k < L; k++)
              if (k == kd) th[k] += A.Lf * Sc / (mu[k] * pib[k]);
            dgl(kd, A.Lf * Sc / dt);
            double ptj = 0;
#pragma unroll
            for (int j = 0; j < L; j++) {
              const double pbj = ptj + P.grav * mu[j];
              if (j > kd && Sc > 0 && base_T<L, MIX>(j, th[j], pbj, P, A) > A.Tf) {
                th[j] -= A.Lf * Sc / (mu[j] * pib[j]);
                dgl(j, -(A.Lf * Sc / dt));
                Sc = 0;
              }
              ptj = pbj;
            }
          }
          if (!cu) {
#pragma unroll
            for (int j = 0; j < L - 1; j++) {
              if (j >= kd) wxo[j] += mc;
              if (j >= kd && j < bot - 1) wxo[j] += me;
            }
          }
          cin(kd, mc + me);
          if constexpr (CLOUD) {
#pragma unroll
            for (int k = 0; k < L; k++)
              if (k == kd) cl[k] += mc * det / mu[k];
            cloud_adjust_k<L, MIX>(kd, mu, th, q, cl, pib, dt, P, A);
          } else {
            const double rt = [&] {
              double r = 0;
#pragma unroll
              for (int k = 0; k < L - 1; k++)
                if (k == kd) r = mu[k] * condense_k<L, MIX>(k, mu, th, q, pib, P, A);
              return r;
            }();
            R += rt;
            Rc += rt;
          }
          if (A.downdraft > 0 && bot >= 2 && Rc - Sc > 0) {
            // the downdraft from the layer above the bottom one; with cu_dth from the cumulus layer or the cloud layer
            // above it, whichever has the lower equivalent potential theta exp(Lv q / (cp T)) at its middle
            auto downdraft = [&](auto S_) {
              constexpr int s_ = decltype(S_)::value;
              double pt_ = 0;
#pragma unroll
              for (int k = 0; k < s_; k++) pt_ += P.grav * mu[k];
              const double p1 = pt_ + 0.5 * P.grav * mu[s_];
              const double p2 = s_ == bot - 1 ? pt_ + P.grav * (mu[s_] + mu[bot])
                                              : pt_ + P.grav * (mu[s_] + mu[s_ + 1 < L ? s_ + 1 : bot] + mu[bot]);
              double thd_ = th[s_], qd_ = q[s_];
              for (int j = 0; j <= 4; j++) wet_bulb<MIX>(thd_, qd_, p1 + (p2 - p1) * j / 4, P, A);
              const double ev = qd_ - q[s_];
              double md = fmin(A.downdraft * mc, fmax(0.0, mu[s_] - P.floor_));
              if (ev > 0) md = fmin(md, (Rc - Sc) / ev);
              if (md > 0 && thd_ < th[s_]) {
                move_mass<L, true, CLOUD>(P, mu, th, q, pib, s_, bot, md, th[s_], thd_, q[s_], qd_, cl);
#pragma unroll
                for (int j = 0; j < L - 1; j++)
                  if (j >= s_) wxo[j] -= md;
                const double e_ = md * fmax(0.0, ev);
                R -= e_;
                Rc -= e_;
                dgl(bot, -(P.Lv * e_ / dt));
              }
            };
            bool lower = false;
            if constexpr (L >= 4) {
              if (cu) {
                double pc = 0, te_[2];
#pragma unroll
                for (int k = 0; k < bot - 2; k++) pc += P.grav * mu[k];
#pragma unroll
                for (int m = 0; m < 2; m++) {
                  const int k = bot - 2 + m;
                  const double pm = pc + 0.5 * P.grav * mu[k], Tk = th[k] * fpow_<MIX>(pm / P.p0, P.kappa);
                  te_[m] = th[k] * exp(P.Lv * q[k] / (P.cp * Tk));
                  pc += P.grav * mu[k];
                }
                lower = te_[0] < te_[1];
              }
            }
            if constexpr (L >= 4) {
              if (lower) downdraft(std::integral_constant<int, (L >= 4 ? bot - 2 : bot - 1)>{});
              else downdraft(std::integral_constant<int, bot - 1>{});
            } else {
              downdraft(std::integral_constant<int, bot - 1>{});
            }
          }
          S += Sc;
          Scg = Sc;
        } else {
          mc = 0;
        }
      }
    } else {
    parcel<MIX>(th0, q0, pi_, P, A, thp, qc);
    const double d = thp - tstar<L>(bot - 1, P);
    if (d > 0) {
      const double dp = 500;
      parcel<MIX>(th0, q0, pi_ + dp, P, A, thp2, qc2);
      const double slope = (thp - thp2) / dp;
      const double maxup = mu[bot] - P.floor_;
      const double want = slope > 0 ? d / slope / P.grav : maxup;
      mc = fmin(want, maxup) * omexp<MIX>(dt / A.tau_conv);
      if (mc > 0) {
        int kd = bot - 1;
        double pj = pi_;
        while (kd > 0) {
          double mk = 0;
#pragma unroll
          for (int j = 0; j < L; j++)
            if (j == kd) mk = mu[j];
          pj -= P.grav * mk;
          double t2, q2;
          parcel<MIX>(th0, q0, pj, P, A, t2, q2);
          if (!(t2 > tstar<L>(kd - 1, P))) break;
          thp = t2;
          qc = q2;
          kd--;
        }
        R += mc * (q0 - qc);
        Rc = mc * (q0 - qc);
        move_mass<L, true>(P, mu, th, q, pib, bot, kd, mc, th0, thp, q0, qc);
#pragma unroll
        for (int j = 0; j < L - 1; j++)
          if (j >= kd) wxo[j] += mc;
        cin(kd, mc);
        const double rt = [&] {
          double r = 0;
#pragma unroll
          for (int k = 0; k < L - 1; k++)
            if (k == kd) r = mu[k] * condense_k<L, MIX>(k, mu, th, q, pib, P, A);
          return r;
        }();
        R += rt;
        Rc += rt;
      } else {
        mc = 0;
      }
    }
    }
    // shallow (air.cu:1730-1835, without the cumulus layer): a second plume from the same root (plume_root with
    // root_q), entraining twice the deep one's with ent_ifs, stopping in the cloud layer and leaving its air there;
    // its base mass flux Grant (2001)'s rho a w* (a = shallow), cut to the air and vapor the layers hold. Its
    // condensate leaves as the deep plume's: l_det as cloud, the rest rains, its frozen share as snow
    if (A.shallow > 0 && A.plume_deep > 0 && bot >= 2 && ex_r >= 0 && ws_r > 0) {
      constexpr int ks = bot >= 1 ? bot - 1 : 0;
      double pis = 0;
#pragma unroll
      for (int k = 0; k < bot; k++) pis += P.grav * mu[k];
      const double ths0 = th[bot];
      double qs0 = edge_q<L, MIX>(bot, false, mu, th, q, 0, P, A);
      qs0 += fmin(ex_r, fmax(0.0, qss_r - qs0));
      double As, Gs = 1, FZs = 0, thps, qcs, qxs[L], lts[4], xss = 0, Gls[L];
      int kds;
      plume_ascent<L, MIX>(mu, th, q, ths0, qs0, pis, P, A, As, kds, thps, qcs, qxs, lts, Gs, FZs, Gls, true);
#pragma unroll
      for (int k = 0; k <= bot; k++) xss += qxs[k];
      if (A.cu_dth > 0) {
        // cu_dth: through the cumulus layer and into the cloud layer above it, stopping in either; what it took in
        // from each moved by plume_out; with buoy_v only through the cloud at the mixed layer's top, its share there
        // at most
        if ((kds == bot - 1 || kds == bot - 2) && As > 0) {
          double ms = A.shallow * rho_r * ws_r * dt, cond, det, Gx;
          if (A.buoy_v > 0) {
            double l_, f_, d_;
            ms = mixed_cloud<L, MIX>(bot, mu, th[bot], q[bot] + (CLOUD ? cl[bot] : 0.0), P, A, l_, f_, d_,
                                     A.cu_cloud > 0)
                     ? fmin(ms, rho_r * f_ * ws_r * dt)
                     : 0.0;
          }
          if (qs0 > 0) ms = fmin(ms, mu[bot] * q[bot] / qs0);
          plume_out<L, CLOUD>(P, mu, th, q, pib, cl, kds, ms, ths0, qs0, thps, qcs, Gls, qxs, wxo, A.l_det, cond, det,
                              Gx);
          if (ms > 0) {
            const double rs = ms * (cond - det);
            R += rs;
            Rc += rs;
            if (A.ice_air > 0 && cond > det) {         // (its frozen share snows, melting below where warm)
              double Ss = rs * fmin(1.0, FZs / cond);
#pragma unroll
              for (int k = 0; k < L; k++)
                if (k == kds) th[k] += A.Lf * Ss / (mu[k] * pib[k]);
              dgl(kds, A.Lf * Ss / dt);
              double ptj = 0;
#pragma unroll
              for (int j = 0; j < L; j++) {
                const double pbj = ptj + P.grav * mu[j];
                if (j > kds && Ss > 0 && base_T<L, MIX>(j, th[j], pbj, P, A) > A.Tf) {
                  th[j] -= A.Lf * Ss / (mu[j] * pib[j]);
                  dgl(j, -(A.Lf * Ss / dt));
                  Ss = 0;
                }
                ptj = pbj;
              }
              S += Ss;
              Scg += Ss;
            }
            cin(kds, ms * Gx);
            if constexpr (CLOUD) {
#pragma unroll
              for (int k = 0; k < L; k++)
                if (k == kds) cl[k] += ms * det / mu[k];
              cloud_adjust_k<L, MIX>(kds, mu, th, q, cl, pib, dt, P, A);
            }
            mc += ms;
          }
        }
      } else if (kds == ks && As > 0) {
        // its air taken in from the cloud layer (Gs - 1 per kg leaving the mixed layer) leaves it at the layer's vapor
        // and comes back at the plume's: at most the layer's air above its floor, and what it and the mixing take at
        // most half the layer's vapor
        const double qe = q[ks];
        double ms = fmin(A.shallow * rho_r * ws_r * dt, mu[bot] - P.floor_);
        if (qs0 > 0) ms = fmin(ms, mu[bot] * q[bot] / qs0);
        if (Gs > 1) ms = fmin(ms, (mu[ks] - P.floor_) / (Gs - 1));
#pragma unroll
        for (int k = 0; k < bot; k++) {
          const double take = (k == ks ? (Gs - 1) * qe : 0.0) - qxs[k];
          if (take > 0) ms = fmin(ms, 0.5 * mu[k] * q[k] / take);
        }
        if (ms > 0) {
          const double cond_ = qs0 + (Gs - 1) * qe - Gs * qcs - xss, cond = fmax(0.0, cond_);
          const double det = CLOUD ? fmin(cond, A.l_det * Gs) : 0.0, rs = ms * (cond - det);
          const double qarr = Gs * qcs + (cond_ - cond) - (Gs - 1) * qe;   // (per kg moved: Gs kg arrive)
          const double Ec = column_energy<L, MOIST>(mu, pib, th, q, P);
#pragma unroll
          for (int k = 0; k < bot; k++) q[k] = fmax(0.0, q[k] + ms * qxs[k] / mu[k]);
          move_mass<L, true, CLOUD>(P, mu, th, q, pib, bot, ks, ms, ths0, thps, qs0, qarr, cl);
          th[ks] += (Ec - column_energy<L, MOIST>(mu, pib, th, q, P)) / (mu[ks] * pib[ks]);
          R += rs;
          Rc += rs;
          if (A.ice_air > 0 && cond > det) {           // (its frozen share snows, melting in a warm bottom layer)
            const double Ss = rs * fmin(1.0, FZs / cond);
            th[ks] += A.Lf * Ss / (mu[ks] * pib[ks]);
            dgl(ks, A.Lf * Ss / dt);
            double ps_ = 0;
#pragma unroll
            for (int k = 0; k < L; k++) ps_ += P.grav * mu[k];
            if (base_T<L, MIX>(bot, th[bot], ps_, P, A) > A.Tf) {
              th[bot] -= A.Lf * Ss / (mu[bot] * pib[bot]);
              dgl(bot, -(A.Lf * Ss / dt));
            } else {
              S += Ss;
              Scg += Ss;
            }
          }
          wxo[ks] += ms;
          cin(ks, ms);
          if constexpr (CLOUD) {
            cl[ks] += ms * det / mu[ks];
            cloud_adjust_k<L, MIX>(ks, mu, th, q, cl, pib, dt, P, A);
          }
          mc += ms;
        }
      }
    }
    // cu_sub (air.cu:1834-1861): the air the plumes took out of the mixed layer comes back from the cumulus layer,
    // sinking between the clouds, up to the condensation level: the slice from that layer's bottom edge, its theta
    // and vapor as the mixed layer's entrainment's; sub_mix stirs it toward the cumulus layer's mean
    if constexpr (L >= 3) {
      if (A.cu_sub > 0 && A.bl_lcl > 0 && blm) {
        constexpr int ka = bot - 1;
        const double m = fmin(fmin(mu_pre - mu[bot], lcl_mass<L, MIX>(mu, th[bot], q[bot], P, A) - mu[bot]),
                              mu[ka] - P.floor_);
        if (m > 0) {
          const double e = tstar<L>(ka, P), fm = fmin(1.0, m / mu[ka]);
          const bool spread = th[ka] < tstar<L>(ka - 1, P) && th[ka] > e;
          double thc = spread ? e + fm * (th[ka] - e) : th[ka];
          double qx = A.edge_vap > 0 ? edge_q<L, MIX>(ka, false, mu, th, q, m, P, A) : q[ka];
          if (A.sub_mix > 0) {
            thc += A.sub_mix * (th[ka] - thc);
            qx += A.sub_mix * (q[ka] - qx);
          }
          move_mass<L, true, CLOUD>(P, mu, th, q, pib, ka, bot, m, thc, thc, qx, qx, cl);
          wxo[ka] -= m;
        }
      }
    }
  }
  // the plumes' rain evaporates as it falls below their base through the bottom layer's clear air (revap: on the
  // convection's area revap_area; else on conv_revap of the cell; with downdrafts they evaporate it instead)
  if constexpr (CLOUD) {
    if ((A.revap > 0 ? Rc - Scg > 0 : A.conv_revap > 0 && !(A.downdraft > 0) && Rc > 0) && L > 1) {
      double pt_ = 0;
#pragma unroll
      for (int k = 0; k < bot; k++) pt_ += P.grav * mu[k];
      F qs, dqs;
      layer_cap_k<L, MIX>(bot, th[bot], pt_, pt_ + P.grav * mu[bot], P, A, qs, dqs);
      const double rc_ = A.revap > 0 ? (isl ? A.revap_rhl : A.revap_rhs) : 1.0;
      if (rc_ * qs > q[bot]) {
        double ev;
        const double Fc = A.revap > 0 ? (Rc - Scg) / dt : Rc / dt;
        if (A.revap > 0) {
          const double pm = pt_ + 0.5 * P.grav * mu[bot], ps_ = pt_ + P.grav * mu[bot];
          ev = A.revap_area * 5.44e-4 * (rc_ * qs - q[bot]) * pow(sqrt(pm / ps_) / 5.09e-3 * Fc / A.revap_area, 0.577) *
               mu[bot];
          ev = fmin(ev, fmin(Fc, (rc_ * qs - q[bot]) * mu[bot] / dt / (1 + P.Lv / pib[bot] * dqs)));
        } else {
          ev = A.k_revap * (1 - q[bot] / qs) * sqrt(A.conv_revap * Fc) * mu[bot];
          ev = fmin(ev, fmin(Fc, (qs - q[bot]) * mu[bot] / dt / (1 + P.Lv / pib[bot] * dqs)));
        }
        q[bot] += ev * dt / mu[bot];
        th[bot] -= P.Lv * ev * dt / (mu[bot] * pib[bot]);
        dgl(bot, -(P.Lv * ev));
        R -= ev * dt;
        Rc -= ev * dt;
      }
    }
  }
  if (o.conv) o.conv[i] = mc;
#pragma unroll
  for (int j = 0; j + 1 < L; j++) o.wx[(size_t)j * N + i] = wxo[j];
  if (o.xfer)
    for (int g = 0; g < 3; g++)
#pragma unroll
      for (int k = 0; k < L; k++) o.xfer[((size_t)g * L + k) * N + i] = xfl[g * L + k];
  if (o.dgo) {
#pragma unroll
    for (int k = 0; k < L; k++) o.dgo[(size_t)k * N + i] = Q[k];
  }
  o.rain[i] = R / dt;
  if (!isl && S > 0) {                               // snow into the sea melts with the slab's heat
    Ts[i] -= A.Lf * S / Cw;
    if (o.qocean) o.qocean[i] -= A.Lf * S / dt;
  }
  if (isl) {
    // the ice's melt runs into the ponds, from the snow first. On ice the skin decides: below freezing rain and snow
    // freeze as they land, a pond as fast as the surface can lose the heat (to the cold ice, the cold air, the sky;
    // and into snow without firn at 10 mm/h), the freezing heat into the air touching the ground; at freezing rain
    // runs on and the snow melts from that air. Bare ground: snow lies on frozen ground, its freezing (the rain's
    // share) warming it; else rain runs to the ponds. A pond on frozen ground freezes up to its heat room; on thawed
    // ground it soaks in at k_inf toward the bucket's capacity, and a full bucket springs.
    const double scap = o.soilcap[i];
    pw += ice0 - m_ice;
    double sn = o.snow[i] - (ice0 - m_ice);
    if (skin) {
      if (tsk < A.Tf) {
        const double room = (A.k_ice / dice * (A.Tf - tsk) + hs * fmax(0.0, A.Tf - Tsa) +
                             fmax(0.0, A.eps_s * A.sigma * A.Tf * A.Tf * A.Tf * A.Tf - ldown - swS)) * dt / A.Lf +
                            (sn > 0 && !fon ? 10.0 / 3600 * dt : 0.0);
        const double mp = fmin(pw > 0 ? pw : 0.0, room), m = R + mp;
        m_ice += m;
        sn = (sn > 0 ? sn : 0.0) + R;
        pw -= mp;
#pragma unroll
        for (int k = 0; k <= bot; k++) th[k] += A.Lf * (m - S) * (wbl[k] / mb) / (mu[k] * pib[k]);
      } else {
        pw += R;
#pragma unroll
        for (int k = 0; k <= bot; k++) th[k] -= A.Lf * S * (wbl[k] / mb) / (mu[k] * pib[k]);
      }
      if (fon && pw > 0 && Tfn < A.Tf) {             // melt and rain soak into the cold firn and refreeze
        const double mr = fmin(pw, Cf * (A.Tf - Tfn) / A.Lf);
        pw -= mr;
        m_ice += mr;
        sn = (sn > 0 ? sn : 0.0) + mr;
        Tfn += A.Lf * mr / Cf;
      }
    } else if (Ts[i] < A.Tf) {
      m_ice += R;
      sn = (sn > 0 ? sn : 0.0) + R;
      Ts[i] += A.Lf * (R - S) / Cw;
    } else {
      pw += R;
      Ts[i] -= A.Lf * S / Cw;
    }
    if (!skin && Ts[i] < A.Tf && pw > 0) {
      const double room = (A.Tf - Ts[i]) * Cw / A.Lf, m = pw < room ? pw : room;
      pw -= m;
      m_ice += m;
      Ts[i] += A.Lf * m / Cw;
    }
    o.ice[i] = m_ice;
    o.snow[i] = fmax(0.0, fmin(sn, m_ice));
    if (!skin && Ts[i] >= A.Tf) {
      const double room = A.k_inf * dt * (1 - wet / scap);
      const double m = pw < room ? pw : room;
      if (m > 0) {
        pw -= m;
        wet += m;
      }
    }
    if (wet > scap) {
      pw += wet - scap;
      wet = scap;
    }
    o.soil[i] = wet;
    o.pond[i] = pw;
  }
  if (o.rainp) {
    o.rainp[i] = Rk0 / dt;
    o.rainp[(size_t)N + i] = Rk1 / dt;
    o.rainp[(size_t)2 * N + i] = Rc / dt;
    o.rainp[(size_t)3 * N + i] = S / dt;
  }
  // a layer emptied to its floor keeps a theta within its edges: its heat beyond them goes to the nearest layer with
  // mass (below first), the column's energy unchanged
  if (blm) {
#pragma unroll
    for (int k = 1; k < bot; k++) {
      if (mu[k] > 2 * P.floor_) continue;
      const double lo = tstar<L>(k, P), hi = tstar<L>(k - 1, P), tn = fmin(fmax(th[k], lo), hi);
      if (tn == th[k]) continue;
      int n = -1;
#pragma unroll
      for (int j = k + 1; j < L; j++)
        if (n < 0 && mu[j] > 2 * P.floor_) n = j;
#pragma unroll
      for (int j = k - 1; j >= 0; j--)
        if (n < 0 && mu[j] > 2 * P.floor_) n = j;
      if (n < 0) continue;
      const double moved = mu[k] * pib[k] * (th[k] - tn);
#pragma unroll
      for (int j = 0; j < L; j++)
        if (j == n) th[j] += moved / (mu[j] * pib[j]);
      th[k] = tn;
    }
  }
  const double E1 = column_energy<L, MOIST>(mu, pib, th, q, P) + Cw * Ts[i] - A.Lf * m_ice + Cf * (Tfn - A.Tf);
  if (Cf > 0) o.firn[i] = Tfn;
#pragma unroll
  for (int k = 0; k < L; k++) {
    o.theta[(size_t)k * N + i] = th[k];
    if constexpr (MOIST) o.qv[(size_t)k * N + i] = q[k];
    if constexpr (CLOUD) o.cloudw[(size_t)k * N + i] = cl[k];
  }
  o.stored[i] = E1 - E0;
}

#define ISW_BY_L(CALL)       \
  switch (L) {               \
    case 2: CALL(2); break;  \
    case 3: CALL(3); break;  \
    case 4: CALL(4); break;  \
    case 5: CALL(5); break;  \
    case 6: CALL(6); break;  \
    case 7: CALL(7); break;  \
    case 8: CALL(8); break;  \
    default: return ISW_EBADARG; \
  }

#if ISW_COLUMN_CLOUDS
// The cloudy air column's launches, compiled in a unit of their own (ISW_COLUMN_CLOUDS=1) so that its fourteen
// kernels build in parallel with the rest of this file.
int air_column_cloudy(int N, int L, int mixed, bool full, const CP& P, const AP& A, int mode, const double* t,
                      const double* insol, const double* pos, const double* uabs, const double* mu, double dt,
                      const double* elev, const double* radf, const double* swf, const double* ch, const AirIO& o,
                      double bytes, hipStream_t st) {
#define CALL(LL)                                                                                                     \
  if (mixed) {                                                                                                       \
    if (full) LAUNCHC(LL, true, true);                                                                               \
    else LAUNCHC(LL, true, false);                                                                                   \
  } else {                                                                                                           \
    if (full) LAUNCHC(LL, false, true);                                                                              \
    else LAUNCHC(LL, false, false);                                                                                  \
  }
#define LAUNCHC(LL, MX, FU)                                                                                          \
  ISW_LAUNCH("air_column", bytes, (k_air_column<LL, MX, true, true, FU>), grid1(N), dim3(BS), N, P, A, mode, t,      \
             insol, pos, uabs, mu, dt, elev, radf, swf, ch, o)
  ISW_BY_L(CALL)
#undef CALL
#undef LAUNCHC
  return (int)hipGetLastError();
}
#else
// (defined in the ISW_COLUMN_CLOUDS=1 unit)
int air_column_cloudy(int N, int L, int mixed, bool full, const CP& P, const AP& A, int mode, const double* t,
                      const double* insol, const double* pos, const double* uabs, const double* mu, double dt,
                      const double* elev, const double* radf, const double* swf, const double* ch, const AirIO& o,
                      double bytes, hipStream_t st);

int hs_column(int N, int L, const double* par, const double* hs, const double* lat, const double* mu,
              double* theta, double dt, double* wx, double* qnet, double* stored, hipStream_t st) {
  const CP P = make_cp(par);
  const double D = 8;
  const double bytes = N * D * (1 + L + 2.0 * L + (L - 1) + 2);
#define CALL(LL) \
  ISW_LAUNCH("hs_column", bytes, k_hs_column<LL>, grid1(N), dim3(BS), N, P, hs[0], hs[1], lat, mu, theta, dt, wx, qnet, stored)
  ISW_BY_L(CALL)
#undef CALL
  return (int)hipGetLastError();
}

int surface_layer(int N, int L, const double* par, const isw_air* air, const double* uabs, const double* mu,
                  const double* theta, const double* qv, const double* Ts, const double* ice, const double* land,
                  double* cd, double* ch, const double* lc, hipStream_t st) {
  const CP P = make_cp(par);
  const double bytes = 8.0 * N * (1 + L + 1 + (qv ? 1 : 0) + 1 + (ice ? 1 : 0) + (land ? 1 : 0) + (lc ? 1 : 0) + 2);
  ISW_LAUNCH("surface_layer", bytes, k_surface_layer, grid1(N), dim3(BS), N, L, P, air->rd, air->vf, air->eps, uabs,
             mu, theta, qv, Ts, ice, land, cd, ch, lc);
  return (int)hipGetLastError();
}

int air_column(int N, int L, const double* par, const isw_air* air, int mode, const double* t, const double* insol,
               const double* pos, const double* uabs, const double* mu, double* theta, double* qv, double* Ts,
               double* ice, double dt, double* wx, double* olr, double* sw_abs, double* stored, double* rain,
               double* evap, double* conv, double* tsurf, double* rainp, double* cap, double* qocean,
               double* qsw, const double* land, const double* soilcap, double* soil, double* pond, double* snow,
               const double* elev, double* flash, const double* radf, const double* swf, double* swalb,
               const double* ch, const double* lc, double* firn, double* cloudw, double* cloudo, double* xfer,
               double* dgo, int mixed, hipStream_t st) {
  const CP P = make_cp(par);
  const AP A = make_ap(par, air);
  const bool moist = P.moist > 0, cloudy = moist && par[55] > 0 && cloudw;
  if (!moist) land = nullptr;                          // (the dry column has no land)
  const AirIO o{theta, qv, Ts, ice, wx, olr, sw_abs, stored, rain, evap, conv, tsurf, rainp, qocean, qsw, flash,
                swalb, firn, cloudw, cloudo, cap, land, soilcap, lc, soil, pond, snow, xfer, dgo};
  const int RW = L + 1;
  const double rows_in = 1 + 2.0 * L + 1 + (mode == 0 ? 3 : 1) + (cap ? 1 : 0) + (elev ? 1 : 0) + (ch ? 1 : 0) +
                         (radf ? 3.0 * RW + 6 * L : 0) + (swf ? 3.0 * L + 1 : 0) + (firn ? 1 : 0) +
                         (moist ? L + 1.0 : 0) + (cloudy ? L : 0) + (land ? 5 : 0) + (land && lc ? 8 : 0);
  const double rows_out = L + 1 + (L - 1) + 6 + (rainp ? 4 : 0) + (qocean ? 1 : 0) + (qsw ? 1 : 0) + (flash ? 1 : 0) +
                          (swalb ? 1 : 0) + (firn ? 1 : 0) + (moist ? L + 1.0 : 0) + (cloudy ? L : 0) +
                          (cloudy && cloudo ? (radf ? 4.0 * L + 3 : L + 3.0) : 0) + (land ? 3 : 0) +
                          (land && cap ? 1 : 0);
  const double bytes = 8.0 * N * (rows_in + rows_out + (xfer ? 6.0 * L : 0) + (dgo ? 5.0 * L - 2 : 0));
  // the moist column's options that take runtime branches of their own: with none of them on it runs the plain
  // kernel (FULL false), their code compiled away (the dry column is always the full one: bl_mix acts there too)
  const bool full = land || P.bl_mix > 0 || par[51] > 0 || par[53] > 0 || par[86] > 0 || par[85] > 0 ||
                    par[83] > 0 || par[88] > 0 || par[47] > 0 || par[89] > 0 || par[111] > 0 || par[114] > 0;
  if (cloudy)
    return air_column_cloudy(N, L, mixed, full, P, A, mode, t, insol, pos, uabs, mu, dt, elev, radf, swf, ch, o,
                             bytes, st);
#define LAUNCH(LL, MX, MO, FU)                                                                                     \
  ISW_LAUNCH("air_column", bytes, (k_air_column<LL, MX, MO, false, FU>), grid1(N), dim3(BS), N, P, A, mode, t,     \
             insol, pos, uabs, mu, dt, elev, radf, swf, ch, o)
#define CALL(LL)                                    \
  if (mixed) {                                      \
    if (!moist) LAUNCH(LL, true, false, true);      \
    else if (full) LAUNCH(LL, true, true, true);    \
    else LAUNCH(LL, true, true, false);             \
  } else {                                          \
    if (!moist) LAUNCH(LL, false, false, true);     \
    else if (full) LAUNCH(LL, false, true, true);   \
    else LAUNCH(LL, false, true, false);            \
  }
  ISW_BY_L(CALL)
#undef CALL
#undef LAUNCH
  return (int)hipGetLastError();
}

int cell_speed(int N, int E, const double* u, const int* cellEdge, const double* le, const double* de,
               const double* areaCell, double* out, hipStream_t st) {
  const double bytes = 8.0 * E * 3 + 4.0 * 6 * N + 8.0 * N * 2;
  ISW_LAUNCH("cell_speed", bytes, k_cell_speed, grid1(N), dim3(BS), N, u, cellEdge, le, de, areaCell, out);
  return (int)hipGetLastError();
}

int cell_speed_perot(int N, int E, const double* u, const int* cellEdge, const double* op, double* out,
                     hipStream_t st) {
  const double bytes = 8.0 * E + 4.0 * 6 * N + 8.0 * 12 * N + 8.0 * N;
  ISW_LAUNCH("cell_speed_perot", bytes, k_cell_speed_perot, grid1(N), dim3(BS), N, u, cellEdge, op, out);
  return (int)hipGetLastError();
}

int gas_src(int N, int L, int G, const double* src, const double* h, double dt, double* x, hipStream_t st) {
  const double bytes = 8.0 * N * (G + 1 + 2.0 * G);
  ISW_LAUNCH("gas_src", bytes, k_gas_src, grid1(N), dim3(BS), N, L, G, src, h, dt, x);
  return (int)hipGetLastError();
}

int advance_time(double* t, double dt, hipStream_t st) {
  ISW_LAUNCH("advance_time", 16.0, k_advance_time, dim3(1), dim3(1), t, dt);
  return (int)hipGetLastError();
}

#endif  // ISW_COLUMN_CLOUDS
}  // namespace isw_col
