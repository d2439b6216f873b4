// The C ABI of libinsilica_hip.so (include/insilica_hip.h): argument checks on the host, then dispatch to the plain
// FP64 (isw_f64) or the mixed (isw_mix) instantiation of core.hip; and the deterministic sums, which are FP64 in
// both forms.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "diag.hpp"
#include "insilica_hip.h"

#define ISW_DECLARE(NS)                                                                                              \
  namespace NS {                                                                                                     \
  int layers_step(const isw_grid*, const isw_layers*, double*, double*, double*, double*, double, int, hipStream_t); \
  int layers_lap(const isw_grid*, const isw_layers*, const double*, double*, hipStream_t);                          \
  int layers_gfast(const isw_grid*, const isw_layers*, const double*, const double*, hipStream_t);                  \
  int layers_exchange(const isw_grid*, const isw_layers*, const double*, double*, double*, double*, double*,        \
                      double*, int, double, hipStream_t);                                                            \
  int vmix_heat(const isw_grid*, const isw_layers*, const double*, const double*, double*, hipStream_t);            \
  int mass_update(const isw_grid*, int, const double*, const double*, double*, hipStream_t);                        \
  int transport_step(const isw_grid*, const isw_transport*, const double*, const double*, const double*, double*,   \
                     int, int, hipStream_t);                                                                         \
  int swe_step(const isw_grid*, const isw_layers*, double*, double*, double*, double, hipStream_t);                 \
  int couple_stress(const isw_grid*, int, const double*, double*, const double*, double*, double, double, double,   \
                    double*, double*, const double*, double, const double*, hipStream_t);                           \
  int land_column(const isw_grid*, int, const double*, double*, const double*, double*, const double*, double,      \
                  double, double, hipStream_t);                                                                      \
  int ocn_cap(const isw_grid*, const double*, double, double*, const double*, double, hipStream_t);                 \
  int fill(double*, long, double, hipStream_t);                                                                      \
  }
ISW_DECLARE(isw_f64)
ISW_DECLARE(isw_mix)

namespace isw_col {
int hs_column(int, int, const double*, const double*, const double*, const double*, double*, double, double*, double*,
              double*, hipStream_t);
int surface_layer(int, int, const double*, const isw_air*, const double*, const double*, const double*, const double*,
                  const double*, const double*, const double*, double*, double*, const double*, hipStream_t);
int cell_speed(int, int, const double*, const int*, const double*, const double*, const double*, double*, hipStream_t);
int cell_speed_perot(int, int, const double*, const int*, const double*, double*, hipStream_t);
int gas_src(int, int, int, const double*, const double*, double, double*, hipStream_t);
int advance_time(double*, double, hipStream_t);
int air_column(int, int, const double*, const isw_air*, int, const double*, const double*, const double*,
               const double*, const double*, double*, double*, double*, double*, double, double*, double*, double*,
               double*, double*, double*, double*, double*, double*, double*, double*, double*, const double*,
               const double*, double*, double*, double*, const double*, double*, const double*, const double*,
               double*, const double*, const double*, double*, double*, double*, double*, double*, int,
               hipStream_t);
}  // namespace isw_col

namespace isw_land {
int route(const isw_grid*, const double*, const double*, const double*, double*, double*, const double*, double, double,
          double, double, double*, double*, double*, hipStream_t);
int ice(const isw_grid*, const double*, const double*, double*, double, double, double*, double*, double*, double*,
        hipStream_t);
}  // namespace isw_land

namespace isw_ocn {
int overturn(int, int, const isw_ocean*, const double*, double*, double*, const double*, hipStream_t);
int column(int, int, const isw_ocean*, const double*, double*, double*, const double*, const double*, const double*,
           const double*, double, double*, const double*, const double*, const double*, const double*, hipStream_t);
int xch_salt(int, int, const double*, const double*, double*, int, hipStream_t);
int pen(int, int, const isw_ocean*, const double*, double*, const double*, const double*, double, double*, double*,
        const double*, const double*, hipStream_t);
int fresh(int, const isw_ocean*, const double*, double*, const double*, const double*, const double*, double*, double,
          double*, const double*, double*, double*, double*, const double*, hipStream_t);
int diffuse(const isw_grid*, int, const double*, const double*, const double*, double*, double*, const double*, double,
            double, hipStream_t);
int forcing(const isw_grid*, double, double*, double*, double*, double*, double*, hipStream_t);
}  // namespace isw_ocn

#define ISW_DECLARE_RAD(NS)                                                                                          \
  namespace NS {                                                                                                     \
  int lw_setup(int, int, const double*, const isw_lw_tables*, const isw_air*, double, double, double, const double*, \
               const double*, const double*, const double*, const double*, double, int, float*, hipStream_t);         \
  int lw_column(int, int, const double*, const isw_lw_tables*, const float*, const double*, double*, hipStream_t);   \
  int sw_column(int, int, const double*, const isw_sw_tables*, const isw_air*, const float*, int, const double*,     \
                double, const double*, const double*, const double*, double*, hipStream_t);                         \
  }
ISW_DECLARE_RAD(isw_f64)
ISW_DECLARE_RAD(isw_mix)

namespace {

inline hipStream_t S(void* s) { return reinterpret_cast<hipStream_t>(s); }

bool grid_ok(const isw_grid* g) {
  if (!(g && g->N > 0 && g->E > 0 && g->NV > 0 && g->eI && g->eJ && g->cellEdge && g->cellEdgeAsc && g->cellSign &&
        g->le && g->de && g->areaCell))
    return false;
  // the fast sub-steps' patches, when given: every table, a patch count that covers the cells, slots for a patch
  if (g->ph_off && !(g->ph_cell && g->pe_off && g->pe_edge && g->pe_slot && g->eslot && g->pcells > 0 &&
                     g->npatch == (g->N + g->pcells - 1) / g->pcells && g->pslots >= g->pcells &&
                     g->pslots < (1 << 15) && g->pown > 0))
    return false;
  // the transport's fused stages: their incoming edges, each cell's slots, the most edges a patch uses
  if (g->pi_off && !(g->ph_off && g->pi_edge && g->tc_slot && g->ptedges > 0 && g->ptedges < (1 << 14)))
    return false;
  // the pair kernel's patches: one thread of its ISW_PAIR_THREADS-thread block per owned edge, and per cell of patch
  // and halo (at most pslots of them)
  return !g->ph2_off || (g->ph_off && g->ph2_cell && g->px_off && g->px_edge && g->px_slot && g->pc_off &&
                         g->pc_eslot && g->pslots2 >= g->pslots && g->pslots2 < (1 << 15) && g->pedges2 > 0 &&
                         g->pedges2 < 0xffff && g->pown <= ISW_PAIR_THREADS && g->pslots <= ISW_PAIR_THREADS);
}

bool layers_ok(const isw_grid* g, const isw_layers* s) {
  if (!grid_ok(g) || !s || s->L < 1 || s->L > ISW_LMAX) return false;
  if (s->prec != ISW_FP64 && s->prec != ISW_MIXED) return false;
  if (s->prec == ISW_MIXED && (!s->pkf || !g->wf || !g->rA || !g->rde || !g->rAV)) return false;
  // (the patch kernels read the stage's coefficients per cell slot)
  if (s->air && g->ph_off && (s->prec == ISW_MIXED ? !s->cfcf : !s->cfc)) return false;
  if (s->prec == ISW_FP64 && !s->pk) return false;
  const isw_batch* b = s->batch;
  if (b && (b->nworlds < 1 || !b->wc || !b->we || !b->nsub || !b->grav || !b->nu4 || !b->tau || !b->inv_n ||
            b->nmax[0] < 1 || b->nmax[1] < 1 || b->nmax[2] < 1))
    return false;
  return s->fV && s->b && s->Gc && s->B && s->q && s->Tp && s->G && s->He && s->ubs && s->scale && s->h1 && s->h2 &&
         s->u1 && s->u2 && s->D1 && s->Z1 && s->L1 && s->L2 && s->hedge && s->diss && s->hold &&
         (!s->air || (s->pibar && s->S && s->muf && s->mup && s->uf && s->ufb && s->accL)) &&
         (s->air || (s->ub0 && s->ub && s->ubb && s->acc && s->eta0 && s->eta && s->eta_prev));
}

// The radiation's tables: shapes within the kernels' limits (g-points one block, table axes at least two points).
bool lw_tables_ok(const isw_lw_tables* t) {
  return t && t->NC == 6 && t->NP >= 2 && t->NT >= 2 && t->NG >= 1 && t->NG <= 256 && t->NTP >= 2 && t->NO3 >= 2 &&
         t->logk && t->planck && t->o3 && t->dlp != 0 && t->dtt != 0 && t->o3dlp != 0;
}
bool sw_tables_ok(const isw_sw_tables* t) {
  return t && t->NC == 7 && t->NP >= 2 && t->NT >= 2 && t->NG >= 1 && t->NG <= 256 && t->logk && t->sun && t->ray &&
         t->hzf && t->cssa && (!t->cssa_i) == (!t->g_i);
}
bool prec_ok(int32_t p) { return p == ISW_FP64 || p == ISW_MIXED; }

// The air column's options (par, the reference's PAR_ORDER) that this library's column does not carry: none now
// (the cumulus layer's options cu_dth, cu_dp, cl_float, cu_sub and sub_mix are the reference's); kept as the one place
// the entry point and isw_air_column_options ask.
bool column_options_ok(int32_t L, const double* par) {
  (void)L;
  (void)par;
  return true;
}

}  // namespace

#define DISPATCH(prec, call) ((prec) == ISW_MIXED ? isw_mix::call : isw_f64::call)

extern "C" {

int32_t isw_layers_step(const isw_grid* g, const isw_layers* lay, double* h, double* u, double* theta, double* Fm,
                        double dt, int32_t nfast, void* stream) {
  if (!layers_ok(g, lay) || !h || !u || nfast < 1 || !(dt > 0) || (lay->air && !theta)) return ISW_EBADARG;
  return DISPATCH(lay->prec, layers_step(g, lay, h, u, lay->air ? theta : nullptr, Fm, dt, nfast, S(stream)));
}

int32_t isw_layers_lap(const isw_grid* g, const isw_layers* lay, const double* u, double* out, void* stream) {
  if (!layers_ok(g, lay) || !u || !out) return ISW_EBADARG;
  return DISPATCH(lay->prec, layers_lap(g, lay, u, out, S(stream)));
}

int32_t isw_layers_gfast(const isw_grid* g, const isw_layers* lay, const double* h, const double* theta,
                         void* stream) {
  if (!layers_ok(g, lay) || !h || (lay->air && !theta)) return ISW_EBADARG;
  return DISPATCH(lay->prec, layers_gfast(g, lay, h, theta, S(stream)));
}

int32_t isw_layers_exchange(const isw_grid* g, const isw_layers* lay, const double* wx, double* h, double* u,
                            double* M, double* psi, double* theta, int32_t seq, double nu4dt, void* stream) {
  if (!layers_ok(g, lay) || !wx || !h || !u || !M || !psi) return ISW_EBADARG;
  return DISPATCH(lay->prec, layers_exchange(g, lay, wx, h, u, M, psi, theta, seq, nu4dt, S(stream)));
}

int32_t isw_vmix_heat(const isw_grid* g, const isw_layers* lay, const double* h, const double* diss, double* theta,
                      void* stream) {
  if (!layers_ok(g, lay) || !h || !diss || !theta) return ISW_EBADARG;
  return DISPATCH(lay->prec, vmix_heat(g, lay, h, diss, theta, S(stream)));
}

int32_t isw_mass_update(const isw_grid* g, int32_t L, const double* M0, const double* Fm, double* M1, void* stream) {
  if (!grid_ok(g) || L < 1 || L > ISW_LMAX || !M0 || !Fm || !M1) return ISW_EBADARG;
  return isw_f64::mass_update(g, L, M0, Fm, M1, S(stream));
}

int32_t isw_transport_step(const isw_grid* g, const isw_transport* tr, const double* M0, const double* M1,
                           const double* Fm, double* psi, int32_t rows, int32_t order, void* stream) {
  if (!grid_ok(g) || !tr || tr->L < 1 || rows < 1 || rows > tr->rows_max || rows % tr->L || !M0 || !M1 || !Fm ||
      !psi || (order != 1 && order != 3) || !g->nbr || !g->d2idx || !tr->td)
    return ISW_EBADARG;
  if (order == 3 && (!tr->p1 || !tr->p2 || !tr->Rp || !tr->Rm || !tr->Hf || !tr->A)) return ISW_EBADARG;
  if (tr->prec == ISW_MIXED && (!tr->pf || !g->d2wtf)) return ISW_EBADARG;
  if (tr->prec == ISW_FP64 && !g->d2wt) return ISW_EBADARG;
  return DISPATCH(tr->prec, transport_step(g, tr, M0, M1, Fm, psi, rows, order, S(stream)));
}

int32_t isw_swe_step(const isw_grid* g, const isw_layers* lay, double* h, double* u, double* Fm, double dt,
                     void* stream) {
  if (!grid_ok(g) || !lay || lay->L < 1 || lay->L > ISW_LMAX || !h || !u || !lay->F || !lay->B || !lay->q ||
      !lay->scale || !lay->h1 || !lay->h2 || !lay->u1 || !lay->u2 || !lay->b || !lay->fV || !g->eoe || !g->w ||
      (lay->nu4 != 0 && (!lay->D1 || !lay->Z1 || !lay->L1 || !lay->L2)))
    return ISW_EBADARG;
  if (lay->prec == ISW_MIXED && !g->rAV) return ISW_EBADARG;
  return DISPATCH(lay->prec, swe_step(g, lay, h, u, Fm, dt, S(stream)));
}

int32_t isw_couple_stress(const isw_grid* g, int32_t La, const double* ha, double* ua, const double* ho, double* uo,
                          double c, double rho_w, double dt, double* diss, double* tau_acc, const double* wall,
                          double land_drag, const double* cd, void* stream) {
  if (!grid_ok(g) || La < 1 || La > ISW_LMAX || !ha || !ua || !ho || !uo || !diss || !tau_acc || !(dt > 0))
    return ISW_EBADARG;
  return isw_f64::couple_stress(g, La, ha, ua, ho, uo, c, rho_w, dt, diss, tau_acc, wall, land_drag, cd, S(stream));
}

int32_t isw_land_column(const isw_grid* g, int32_t K, const double* land, double* T, const double* cap, double* Tg,
                        const double* dz, double c_soil, double k_soil, double dt, void* stream) {
  if (!g || g->N < 1 || K < 1 || K > ISW_KSOIL || !land || !T || !cap || !Tg || !dz) return ISW_EBADARG;
  return isw_f64::land_column(g, K, land, T, cap, Tg, dz, c_soil, k_soil, dt, S(stream));
}

int32_t isw_ocn_cap(const isw_grid* g, const double* h, double rho_cw, double* cap, const double* land, double c_land,
                    void* stream) {
  if (!g || g->N < 1 || !h || !cap) return ISW_EBADARG;
  return isw_f64::ocn_cap(g, h, rho_cw, cap, land, c_land, S(stream));
}

// ---- the ocean's column processes (FP64 in both forms) ----
bool ocean_ok(const isw_ocean* o) {
  return o && o->grav > 0 && o->rho_w > 0 && o->cw > 0 && o->hmin > 0 && o->hmax >= o->hmin && o->sw_z1 > 0 &&
         o->sw_z2 > 0;
}

int32_t isw_ocn_overturn(int32_t N, int32_t L, const isw_ocean* o, const double* h, double* T, double* S,
                         const double* land, void* stream) {
  if (N < 1 || L < 2 || L > ISW_LMAX || !ocean_ok(o) || !h || !T || !S) return ISW_EBADARG;
  return isw_ocn::overturn(N, L, o, h, T, S, land, reinterpret_cast<hipStream_t>(stream));
}

int32_t isw_ocn_forcing(const isw_grid* g, double dt, double* tau_acc, double* taum, double* q_acc, double* sun_acc,
                        double* fw_acc, void* stream) {
  if (!g || g->N < 1 || g->E < 1 || !g->cellEdge || !g->le || !g->de || !g->areaCell || !(dt > 0) || !tau_acc ||
      !taum || !q_acc || !sun_acc || !fw_acc)
    return ISW_EBADARG;
  return isw_ocn::forcing(g, dt, tau_acc, taum, q_acc, sun_acc, fw_acc, reinterpret_cast<hipStream_t>(stream));
}

int32_t isw_ocn_column(int32_t N, int32_t L, const isw_ocean* o, const double* h, double* T, double* S,
                       const double* taum, const double* qns, const double* qfw, const double* qsun, double dt,
                       double* wx, const double* land, const double* shear, const double* Tc, const double* Sc,
                       void* stream) {
  if (N < 1 || L < 2 || L > ISW_LMAX || !ocean_ok(o) || !h || !T || !S || !taum || !qns || !qfw || !qsun || !wx ||
      !(dt > 0) || (o->det_class > 0 && !Tc))
    return ISW_EBADARG;
  return isw_ocn::column(N, L, o, h, T, S, taum, qns, qfw, qsun, dt, wx, land, shear, Tc, Sc, reinterpret_cast<hipStream_t>(stream));
}

int32_t isw_ocn_xch_salt(int32_t N, int32_t L, const double* wx, const double* h, double* S, int32_t seq,
                         void* stream) {
  if (N < 1 || L < 2 || L > ISW_LMAX || !wx || !h || !S) return ISW_EBADARG;
  return isw_ocn::xch_salt(N, L, wx, h, S, seq, reinterpret_cast<hipStream_t>(stream));
}

int32_t isw_ocn_pen(int32_t N, int32_t L, const isw_ocean* o, const double* h, double* T, const double* qocean,
                    const double* qsw, double dt, double* qns_acc, double* sun_acc, const double* land,
                    const double* up, void* stream) {
  if (N < 1 || L < 2 || L > ISW_LMAX || !ocean_ok(o) || !h || !T || !qocean || !qsw || !qns_acc || !sun_acc)
    return ISW_EBADARG;
  return isw_ocn::pen(N, L, o, h, T, qocean, qsw, dt, qns_acc, sun_acc, land, up, reinterpret_cast<hipStream_t>(stream));
}

int32_t isw_ocn_fresh(int32_t N, const isw_ocean* o, const double* h, double* S, const double* evap,
                      const double* rain, const double* ice, double* ice_prev, double dt, double* fw_acc,
                      const double* land, double* mouth, double* hw, double* M, const double* area, void* stream) {
  if (N < 1 || !ocean_ok(o) || !h || !S || !evap || !rain || !ice || !ice_prev || !fw_acc ||
      (o->fw_local > 0 && (!hw || !M || !area)))
    return ISW_EBADARG;
  return isw_ocn::fresh(N, o, h, S, evap, rain, ice, ice_prev, dt, fw_acc, land, mouth, hw, M, area, reinterpret_cast<hipStream_t>(stream));
}

int32_t isw_ocn_diffuse(const isw_grid* g, int32_t L, const double* h, const double* T, const double* S, double* Tn,
                        double* Sn, const double* wall, double kappa_h, double dt, void* stream) {
  if (!g || g->N < 1 || !g->cellEdge || !g->eI || !g->eJ || !g->le || !g->de || !g->areaCell || L < 1 ||
      L > ISW_LMAX || !h || !T || !S || !Tn || !Sn)
    return ISW_EBADARG;
  return isw_ocn::diffuse(g, L, h, T, S, Tn, Sn, wall, kappa_h, dt, reinterpret_cast<hipStream_t>(stream));
}

int32_t isw_fill(double* x, int64_t n, double v, void* stream) {
  if (n < 0 || (n > 0 && !x)) return ISW_EBADARG;
  return isw_f64::fill(x, (long)n, v, S(stream));
}

int32_t isw_lw_setup(int32_t N, int32_t L, const double* par, const isw_lw_tables* t, const isw_air* air, double co2,
                     double n2o, double ch4, const double* gx, const double* mu, const double* theta, const double* qv,
                     const double* cloudo, double hs, int32_t tmode, float* sub, int32_t prec, void* stream) {
  if (N < 1 || L < 1 || L > ISW_LMAX || !par || !lw_tables_ok(t) || !air || !mu || !theta || !sub || !(hs > 0) ||
      tmode < 0 || tmode > 2 || !prec_ok(prec) || (air->cia && !t->cia))
    return ISW_EBADARG;
  return DISPATCH(prec, lw_setup(N, L, par, t, air, co2, n2o, ch4, gx, mu, theta, qv, cloudo, hs, tmode, sub,
                                 S(stream)));
}

int32_t isw_lw_column(int32_t N, int32_t L, const double* par, const isw_lw_tables* t, const float* sub,
                      const double* Ts, double* radf, int32_t prec, void* stream) {
  if (N < 1 || L < 1 || L > ISW_LMAX || !par || !lw_tables_ok(t) || !sub || !Ts || !radf || !prec_ok(prec))
    return ISW_EBADARG;
  return DISPATCH(prec, lw_column(N, L, par, t, sub, Ts, radf, S(stream)));
}

int32_t isw_sw_column(int32_t N, int32_t L, const double* par, const isw_sw_tables* t, const isw_air* air,
                      const float* sub, int32_t mode, const double* tnow, double interval, const double* pos,
                      const double* alb, const double* cloudo, double* swf, int32_t prec, void* stream) {
  if (N < 1 || L < 1 || L > ISW_LMAX || !par || !sw_tables_ok(t) || !air || !sub || !alb || !swf ||
      (mode != 0 && mode != 1) || (mode == 0 && (!tnow || !pos || !(interval > 0))) || !prec_ok(prec))
    return ISW_EBADARG;
  return DISPATCH(prec, sw_column(N, L, par, t, air, sub, mode, tnow, interval, pos, alb, cloudo, swf, S(stream)));
}

int32_t isw_hs_column(int32_t N, int32_t L, const double* par, const double* hs, const double* lat, const double* mu,
                      double* theta, double dt, double* wx, double* qnet, double* stored, void* stream) {
  if (N < 1 || L < 2 || L > ISW_LMAX || !par || !hs || !lat || !mu || !theta || !wx || !qnet || !stored ||
      !(dt > 0) || (par[70] > 0 && L < 3))
    return ISW_EBADARG;
  return isw_col::hs_column(N, L, par, hs, lat, mu, theta, dt, wx, qnet, stored, S(stream));
}

int32_t isw_air_column(int32_t N, int32_t L, const double* par, const isw_air* air, int32_t mode, const double* t,
                       const double* insol, const double* pos, const double* uabs, const double* mu, double* theta,
                       double* qv, double* Ts, double* ice, double dt, double* wx, double* olr, double* sw_abs,
                       double* stored, double* rain, double* evap, double* conv, double* tsurf, double* rainp,
                       double* dgo, double* cap, double* qocean, double* qsw, const double* land,
                       const double* soilcap, double* soil, double* pond, double* snow, double* cloudw, double* cloudo,
                       const double* elev, double* flash, const double* radf, const double* swf, double* swalb,
                       const double* ch, const double* lc, double* firn, double* xfer, int32_t prec, void* stream) {
  if (N < 1 || L < 2 || L > ISW_LMAX || !par || !air || (mode != 0 && mode != 1) || !uabs || !mu || !theta ||
      !Ts || !wx || !olr || !sw_abs || !stored || !rain || !evap || !conv || !tsurf || !(dt > 0) || !prec_ok(prec))
    return ISW_EBADARG;
  if (mode == 0 ? (!t || !pos) : !insol) return ISW_EBADARG;
  if (!column_options_ok(L, par)) return ISW_EBADARG;
  // moist air needs its vapor and ice; land (moist) its bucket's capacity, soil, ponds and snow
  if (par[22] > 0 && (!qv || !ice || (land && (!soilcap || !soil || !pond || !snow)))) return ISW_EBADARG;
  return isw_col::air_column(N, L, par, air, mode, t, insol, pos, uabs, mu, theta, qv, Ts, ice, dt, wx, olr, sw_abs,
                             stored, rain, evap, conv, tsurf, rainp, cap, qocean, qsw, land, soilcap, soil, pond, snow,
                             elev, flash, radf, swf, swalb, ch, lc, firn, cloudw, cloudo, xfer, dgo,
                             prec == ISW_MIXED, S(stream));
}

int32_t isw_surface_layer(int32_t N, int32_t L, const double* par, const isw_air* air, const double* uabs,
                          const double* mu, const double* theta, const double* qv, const double* Ts, const double* ice,
                          const double* land, double* cd, double* ch, const double* lc, void* stream) {
  if (N < 1 || L < 1 || L > ISW_LMAX || !par || !air || !uabs || !mu || !theta || !Ts || !cd || !ch) return ISW_EBADARG;
  return isw_col::surface_layer(N, L, par, air, uabs, mu, theta, qv, Ts, ice, land, cd, ch, lc, S(stream));
}

int32_t isw_cell_speed(const isw_grid* g, const double* u, double* out, void* stream) {
  if (!grid_ok(g) || !u || !out) return ISW_EBADARG;
  return isw_col::cell_speed(g->N, g->E, u, g->cellEdge, g->le, g->de, g->areaCell, out, S(stream));
}

int32_t isw_cell_speed_perot(const isw_grid* g, const double* u, const double* op, double* out, void* stream) {
  if (!grid_ok(g) || !u || !op || !out) return ISW_EBADARG;
  return isw_col::cell_speed_perot(g->N, g->E, u, g->cellEdge, op, out, S(stream));
}

int32_t isw_gas_src(int32_t N, int32_t L, int32_t G, const double* src, const double* h, double dt, double* x,
                    void* stream) {
  if (N < 1 || L < 1 || L > ISW_LMAX || G < 0 || (G > 0 && (!src || !h || !x))) return ISW_EBADARG;
  return G ? isw_col::gas_src(N, L, G, src, h, dt, x, S(stream)) : 0;
}

int32_t isw_advance_time(double* t, double dt, void* stream) {
  if (!t) return ISW_EBADARG;
  return isw_col::advance_time(t, dt, S(stream));
}

int32_t isw_land_route(const isw_grid* g, const double* b, const double* land, const double* soilcap, double* pond,
                       double* soil, const double* ice, double rho_w, double rho_ice, double u_dt, double kseep_dt,
                       double* work, double* mouth, double* qout, void* stream) {
  if (!grid_ok(g) || !b || !land || !soilcap || !pond || !soil || !ice || !work || !mouth || !(rho_w > 0) ||
      !(rho_ice > 0) || !(u_dt >= 0) || !(kseep_dt >= 0))
    return ISW_EBADARG;
  return isw_land::route(g, b, land, soilcap, pond, soil, ice, rho_w, rho_ice, u_dt, kseep_dt, work, mouth, qout,
                         S(stream));
}

int32_t isw_ice_flow(const isw_grid* g, const double* b, const double* land, double* ice, double rho_ice,
                     double gam_dt, double* work, double* ice_prev, double* calved, double* iout, void* stream) {
  if (!grid_ok(g) || !b || !land || !ice || !work || !calved || !(rho_ice > 0) || !(gam_dt >= 0)) return ISW_EBADARG;
  return isw_land::ice(g, b, land, ice, rho_ice, gam_dt, work, ice_prev, calved, iout, S(stream));
}

const char* isw_version(void) { return "insilica_hip 0.1 gfx950"; }

int32_t isw_grid_check(const isw_grid* g) { return grid_ok(g) ? 0 : ISW_EBADARG; }

int32_t isw_air_column_options(int32_t L, const double* par) {
  if (L < 2 || L > ISW_LMAX || !par) return ISW_EBADARG;
  return column_options_ok(L, par) ? 0 : ISW_EBADARG;
}

}  // extern "C"

// ---------------------------------------------------------------------------------------------------------------
// Deterministic sums (both forms FP64).
// ---------------------------------------------------------------------------------------------------------------
namespace {

constexpr int SUM_THREADS = 256, SUM_CHUNK = 8;

// block_sum (trisk.cu:66-81): thread t adds x[base + t + 256 k], k = 0..7, then a halving tree in LDS.
__global__ __launch_bounds__(SUM_THREADS) void k_block_sum(long long n, const double* __restrict__ x,
                                                           const double* __restrict__ y, double* __restrict__ out) {
  __shared__ double s[SUM_THREADS];
  const long long base = (long long)blockIdx.x * SUM_THREADS * SUM_CHUNK;
  double acc = 0;
  for (int k = 0; k < SUM_CHUNK; k++) {
    const long long i = base + threadIdx.x + (long long)k * SUM_THREADS;
    if (i < n) acc += y ? x[i] * y[i] : x[i];
  }
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int h = SUM_THREADS / 2; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) s[threadIdx.x] += s[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = s[0];
}

// Neumaier's compensated addition of v into (s, c).
__device__ __forceinline__ void neumaier(double& s, double& c, double v) {
  const double t = s + v;
  if (fabs(s) >= fabs(v)) c += (s - t) + v;
  else c += (v - t) + s;
  s = t;
}

// Pass 1: each block's compensated partial (sum, error) over 2048 consecutive values, in a fixed order: thread t
// sums its 8 values, then thread 0 folds the 256 (sum, error) pairs in thread order.
__global__ __launch_bounds__(SUM_THREADS) void k_csum_blocks(long long n, const double* __restrict__ x,
                                                             const double* __restrict__ y, double* __restrict__ part) {
  __shared__ double ss[SUM_THREADS], sc[SUM_THREADS];
  const long long base = (long long)blockIdx.x * SUM_THREADS * SUM_CHUNK;
  double s = 0, c = 0;
  for (int k = 0; k < SUM_CHUNK; k++) {
    const long long i = base + (long long)threadIdx.x * SUM_CHUNK + k;
    if (i < n) neumaier(s, c, y ? x[i] * y[i] : x[i]);
  }
  ss[threadIdx.x] = s;
  sc[threadIdx.x] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    double S0 = 0, C0 = 0;
    for (int t = 0; t < SUM_THREADS; t++) {
      neumaier(S0, C0, ss[t]);
      C0 += sc[t];
    }
    part[2 * blockIdx.x] = S0;
    part[2 * blockIdx.x + 1] = C0;
  }
}

// Pass 2: one block folds the partials in block order, each wave-sized slice by one thread, then thread 0.
__global__ __launch_bounds__(SUM_THREADS) void k_csum_final(long long nb, const double* __restrict__ part,
                                                            double* __restrict__ out) {
  __shared__ double ss[SUM_THREADS], sc[SUM_THREADS];
  const long long per = (nb + SUM_THREADS - 1) / SUM_THREADS;
  double s = 0, c = 0;
  for (long long k = 0; k < per; k++) {
    const long long b = (long long)threadIdx.x * per + k;
    if (b < nb) {
      neumaier(s, c, part[2 * b]);
      c += part[2 * b + 1];
    }
  }
  ss[threadIdx.x] = s;
  sc[threadIdx.x] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    double S0 = 0, C0 = 0;
    for (int t = 0; t < SUM_THREADS; t++) {
      neumaier(S0, C0, ss[t]);
      C0 += sc[t];
    }
    out[0] = S0 + C0;
    out[1] = C0;
  }
}

}  // namespace

namespace {

// The passes of the fixed-order tree (block_sum repeated to one block, gpu.py:111-126): each pass's block sums go to
// one of two halves of ws in turn, the last pass's single sum to out. ws holds at least ceil(n / 2048) + 1 +
// ceil(n / 2048^2) + 1 doubles; every pass writes its values before the next reads them.
int32_t dsum_passes(const double* x, const double* y, long long n, double* ws, double* out, hipStream_t st) {
  const long long per = (long long)SUM_THREADS * SUM_CHUNK;
  long long m = n;
  const double *xx = x, *yy = y;
  double* buf[2] = {ws, ws + (n + per - 1) / per + 1};
  int which = 0;
  for (;;) {
    const long long blocks = (m + per - 1) / per;
    double* dst = blocks == 1 ? out : buf[which];
    k_block_sum<<<dim3((unsigned)blocks), SUM_THREADS, 0, st>>>(m, xx, yy, dst);
    if (blocks == 1) break;
    xx = dst;
    yy = nullptr;
    m = blocks;
    which ^= 1;
  }
  return (int32_t)hipGetLastError();
}

}  // namespace

extern "C" int32_t isw_dsum(const double* x, const double* y, int64_t n, double* scratch, double* out, void* stream) {
  if (n < 1 || !x || !scratch || !out) return ISW_EBADARG;
  return dsum_passes(x, y, n, scratch, out, S(stream));
}

// isw_dsum's workspace as a file-scope device array: each device the library runs on holds its own copy until the
// process ends. Host code never passes this name to a kernel (on the host it is a stand-in, not a device address):
// it uses the address hipGetSymbolAddress gives for the current device, looked up on the first call there.
static __device__ __align__(256) double scratch[ISW_DSUM_NOALLOC_DOUBLES];
static_assert((long long)ISW_DSUM_NOALLOC_MAX_N / (SUM_THREADS * SUM_CHUNK) + 1 +
                      (long long)ISW_DSUM_NOALLOC_MAX_N / ((long long)SUM_THREADS * SUM_CHUNK * SUM_THREADS * SUM_CHUNK) +
                      1 <=
                  ISW_DSUM_NOALLOC_DOUBLES,
              "the static workspace must hold every pass's partials at the largest n");

namespace {
std::mutex scratch_mu;
std::vector<double*> scratch_dev;  // per device: the array's device address, nullptr until looked up
}  // namespace

extern "C" int32_t isw_dsum_noalloc(const double* x, const double* y, int64_t n, double* out, void* stream) {
  if (n < 1 || n > ISW_DSUM_NOALLOC_MAX_N || !x || !out) return ISW_EBADARG;
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return (int32_t)e;
  double* ws = nullptr;
  {
    std::lock_guard<std::mutex> lock(scratch_mu);
    if ((int)scratch_dev.size() <= dev) scratch_dev.resize(dev + 1, nullptr);
    if (!scratch_dev[dev]) {
      void* p = nullptr;
      e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
      if (e != hipSuccess || !p) return e != hipSuccess ? (int32_t)e : (int32_t)hipErrorInvalidSymbol;
      scratch_dev[dev] = static_cast<double*>(p);
    }
    ws = scratch_dev[dev];
  }
  return dsum_passes(x, y, n, ws, out, S(stream));
}

extern "C" int32_t isw_csum(const double* x, const double* y, int64_t n, double* scratch, double* out, void* stream) {
  if (n < 1 || !x || !scratch || !out) return ISW_EBADARG;
  hipStream_t st = S(stream);
  const long long per = (long long)SUM_THREADS * SUM_CHUNK;
  const long long nb = (n + per - 1) / per;
  k_csum_blocks<<<dim3((unsigned)nb), SUM_THREADS, 0, st>>>(n, x, y, scratch);
  k_csum_final<<<dim3(1), SUM_THREADS, 0, st>>>(nb, scratch, out);
  return (int32_t)hipGetLastError();
}

// ---------------------------------------------------------------------------------------------------------------
// Measurement mode (diag.hpp): per-launch event pairs, aggregated per kernel name.
// ---------------------------------------------------------------------------------------------------------------
namespace isw_diag {
namespace {
struct Rec {
  const char* name;
  double bytes;
  hipEvent_t a, b;
};
bool g_on = false;
std::vector<Rec> g_recs;
std::vector<hipEvent_t> g_pool;

hipEvent_t take() {
  if (!g_pool.empty()) {
    hipEvent_t e = g_pool.back();
    g_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

void recycle() {
  for (auto& r : g_recs) {
    if (r.a) g_pool.push_back(r.a);
    if (r.b) g_pool.push_back(r.b);
  }
  g_recs.clear();
}
}  // namespace

bool enabled() { return g_on; }

bool begin(const char* name, double bytes, hipStream_t st, hipEvent_t* a, hipEvent_t* b) {
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(st, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) return false;
  Rec r{name, bytes, take(), take()};
  if (!r.a || !r.b) return false;
  g_recs.push_back(r);
  *a = r.a;
  *b = r.b;
  return true;
}
}  // namespace isw_diag

extern "C" int32_t isw_timing_start(void) {
  isw_diag::recycle();
  isw_diag::g_on = true;
  return 0;
}

extern "C" int32_t isw_timing_stop(char* out, int32_t cap) {
  using namespace isw_diag;
  g_on = false;
  struct Agg {
    long count = 0;
    double ms = 0, bytes = 0;
  };
  std::map<std::string, Agg> agg;
  for (auto& r : g_recs) {
    if (hipEventSynchronize(r.b) != hipSuccess) return -1;
    float ms = 0;
    if (hipEventElapsedTime(&ms, r.a, r.b) != hipSuccess) return -1;
    Agg& a = agg[r.name];
    a.count++;
    a.ms += ms;
    a.bytes += r.bytes;
  }
  recycle();
  std::string text;
  char line[256];
  for (auto& kv : agg) {
    std::snprintf(line, sizeof line, "%s %ld %.6f %.1f\n", kv.first.c_str(), kv.second.count, kv.second.ms,
                  kv.second.bytes);
    text += line;
  }
  if (!out || cap <= (int32_t)text.size()) return -(int32_t)(text.size() + 1);
  std::memcpy(out, text.c_str(), text.size() + 1);
  return (int32_t)text.size();
}
