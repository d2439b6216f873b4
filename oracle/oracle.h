/* CPU oracle of the TRiSK layered core, FCT transport, the one-layer SWE solver and the air-sea coupling of
 * InSilica World V2 -- TEST INFRASTRUCTURE ONLY.
 *
 * This is a scalar C restatement of the reference's double-precision path, kernel by kernel and in the reference's
 * summation order, used by the tests (as the parity checker of the HIP library) and by bench.py's CPU-baseline leg.
 * Nothing in the product links, loads or calls it. Built with gcc -O2 -ffp-contract=off -fopenmp, so every product
 * rounds on its own as the reference's --fmad=false build does; loops over independent cells/edges/vertices are
 * OpenMP-parallel, which changes no result.
 *
 * Arrays are layer-major x[l * n + i], index tables int32 (width 6 per cell, 3 per vertex, pad -1), signs int8.
 */
#ifndef ISW_ORACLE_H
#define ISW_ORACLE_H
#include <stdint.h>

#define OG_GASN 3 /* the gas tracers the air column carries (CO2, N2O, CH4) */
#define OG_LMAX 8

typedef struct og_grid {
  int32_t N, E, NV, n_eoe;
  const int32_t *eI, *eJ, *nbr, *cellEdge, *cellEdgeAsc, *ev1, *ev2, *vEdge, *vCell, *eoeStart, *eoe, *d2idx;
  const int8_t *cellSign, *vEdgeSign, *vSlot;
  const double *le, *de, *areaCell, *areaV, *fV, *kite, *w, *d2wt;
} og_grid;

/* Settings of one stack of layers (reference: src/insilica/layers.py:161-218 Layers.__init__). */
typedef struct og_layers_cfg {
  int32_t L, air;
  double grav, cp, kappa, p0, rd, nu4, floor_, div_damp, thin, shelf, vmix, vmix_cd, thin_wind;
  double gp[OG_LMAX], drag[OG_LMAX], stress[OG_LMAX];
  const double *fV;   /* (NV) the Coriolis parameter with rotation applied */
  const double *b;    /* (N) bottom */
  const double *wall; /* (E) or NULL */
  /* tides (reference: layers.py:57-58): the tide-generating potential per cell (N) subtracted from every Boussinesq
   * layer's Bernoulli function, the bottom layer's quadratic drag coefficient and the internal tide's drag per edge
   * (E, m/s); NULL / 0 without tides */
  const double *tide, *cbot;
  double cd_bot;
} og_layers_cfg;

typedef struct og_layers og_layers; /* opaque: the settings plus preallocated scratch */

og_layers* og_layers_new(const og_grid* g, const og_layers_cfg* cfg);
void og_layers_tides(og_layers* s, const double* tide, const double* cbot, double cd_bot);
void og_layers_free(og_layers* s);
double* og_layers_Gc(og_layers* s);  /* (N) the fast mode's coefficient (grav at creation) */
double* og_layers_L2(og_layers* s);  /* (L,E) del^4 of the last step's start wind */
double* og_layers_diss(og_layers* s); /* (E) */
void og_layers_gfast(og_layers* s, const double* h, const double* theta);
int og_layers_substeps(og_layers* s, const double* h, double dt);
void og_layers_step(og_layers* s, double* h, double* u, double dt, int nfast, double* Fm, double* theta);
void og_layers_lap(og_layers* s, const double* u, double* out);
void og_layers_exchange(og_layers* s, const double* wx, double* h, double* u, double* M, double* psi, double* theta,
                        int seq, double nu4dt);

/* Transport (reference: src/insilica/transport.py) on R = n_tracers * L rows stacked in psi (rows, N). */
void og_mass_update(const og_grid* g, int L, const double* M0, const double* Fm, double* M1);
void og_transport(const og_grid* g, int L, int rows, double beta, const double* M0, const double* M1,
                  const double* Fm, double* psi, int order);

/* SWE.step (reference: src/insilica/swe.py:124-132): L independent layers, b (N), drag (L). */
void og_swe_step(const og_grid* g, int L, double grav, const double* b, const double* drag, double nu4, double* h,
                 double* u, double dt, double* Fm);

/* couple_stress, land_column, ocn_cap, vmix_heat (reference: kernels/ocean.cu:17-42,434-471; layers.cu:599-621). */
void og_couple_stress(const og_grid* g, int La, const double* ha, double* ua, const double* ho, double* uo, double c,
                      double rho_w, double dt, double* diss, double* tau_acc, const double* wall, double land_drag,
                      const double* cd);
void og_land_column(const og_grid* g, int K, const double* land, double* T, const double* cap, double* Tg,
                    const double* dz, double c_soil, double k_soil, double dt);
void og_ocn_cap(const og_grid* g, const double* h, double rho_cw, double* cap, const double* land, double c_land);
void og_vmix_heat(const og_grid* g, int L, const double* h, const double* diss, double grav, double cp, double kappa,
                  double p0, double* theta);
void og_lay_vmix(const og_grid* g, int L, const double* h, double c, double dt, double* u, double* diss, double cd);

/* gpu.dsum's fixed-order tree (reference: kernels/trisk.cu:61-81, gpu.py:111-126). */
double og_dsum(const double* x, const double* y, int64_t n);

/* ---- k-table radiation (reference: src/insilica/radiation.py, kernels/air.cu:2111-2511); radiation.c ----
 * The double-precision path, with the reference's float32 storage points kept: the tables, the sub-levels' 15
 * properties (sub, float32), the clamped table positions (rounded through fminf/fmaxf), and every g-point's
 * contribution, rounded to float32 before the FP64 sum over g-points. par: the reference's PAR_ORDER array. */
typedef struct og_air {
  double m_air, m_air_kg, o2, ray, broad, o3, eps, rd, vf;
  int32_t cia, pad_;
} og_air;

typedef struct og_lwt {
  int32_t NC, NP, NT, NG, NTP, NO3;
  double lp0, dlp, t0, dtt, tp0, o3lp0, o3dlp;
  const float *logk, *planck, *o3, *cia;
} og_lwt;

typedef struct og_swt {
  int32_t NC, NP, NT, NG;
  double yscale, tau550, hssa, hg, inhom;
  const float *logk, *sun, *ray, *hzf, *cssa, *cssa_i, *g_i, *cia;
} og_swt;

void og_lw_setup(int N, int L, const double* par, const og_lwt* t, const og_air* air, double co2, double n2o,
                 double ch4, const double* gx, const double* mu, const double* theta, const double* qv,
                 const double* cloudo, double hs, int tmode, float* sub);
/* nsub sub-levels per layer (the model's 4; the line-by-line columns' test uses 1: each level its own layer). */
void og_lw_column(int N, int L, int nsub, const double* par, const og_lwt* t, const float* sub, const double* Ts,
                  double* radf);
void og_sw_column(int N, int L, int nsub, const double* par, const og_swt* t, const og_air* air, const float* sub,
                  int mode, double tnow, double interval, const double* pos, const double* alb, const double* cloudo,
                  double* swf);
/* The Delta-Eddington layer alone (air.cu:2358-2382), for the closed-form checks. out: Rd, Td, Rs, Ts, Tdd. */
void og_sw_layer(double tau, double ssa, double g, double mu0, double* out);


/* ---- the air column's Held-Suarez path and the surface exchange (reference: kernels/air.cu); column.c ---- */
void og_hs_column(int N, int L, const double* par, const double* hs, const double* lat, const double* mu,
                  double* theta, double dt, double* wx, double* qnet, double* stored);
void og_surface_layer(int N, int L, const double* par, const og_air* air, const double* uabs, const double* mu,
                      const double* theta, const double* qv, const double* Ts, const double* ice, const double* land,
                      double* cd, double* ch, const double* lc);
void og_cell_speed(int N, const double* u, const int32_t* cellEdge, const double* le, const double* de,
                   const double* areaCell, double* out);
void og_cell_speed_perot(int N, const double* u, const int32_t* cellEdge, const double* op, double* out);
void og_gas_src(int N, int L, int G, const double* src, const double* h, double dt, double* x);
/* air_column (air.cu:876-1972): the dry column, and the moist one over the sea without land or the mixed layer; qv
 * (L, N) and ice (N) read and written when moist, the cloud water cloudw (L, N) with cloud > 0 (cloudo ((4L + 3), N)
 * the clouds' cover, effects and optical depths, when given). mix: the reference's mixed build's float32 arithmetic. */
/* The deep (shal = 0) or shallow (shal = 1) plume's ascent alone (column.c plume_ascent, air.cu:540-666): out = (A,
 * kd, thd, qd, G, FZ, lt[0..3], ptop), qx (L) what its mixing gave each layer per kg, Gl (L) what it took in from each.
 * For the tests' closed-form checks. */
void og_plume_ascent(int L, const double* par, const og_air* air, const double* mu, const double* th, const double* q,
                     double th0, double q0, double pi_, int shal, int mix, double* out, double* qx, double* Gl);
void og_air_column(int N, int L, const double* par, const og_air* air, int mode, const double* t, const double* insol,
                   const double* pos, const double* uabs, const double* mu, double* theta, double* qv, double* Ts,
                   double* ice, double dt, double* wx, double* olr, double* sw_abs, double* stored, double* rain,
                   double* evap, double* conv, double* tsurf, double* rainp, double* cap, double* qocean,
                   double* qsw, const double* elev, double* flash, const double* radf, const double* swf,
                   double* swalb, const double* ch, double* firn, double* cloudw, double* cloudo, const double* land,
                   const double* soilcap, double* soil, double* pond, double* snow, const double* lc, double* xfer,
                   double* dgo, int mix);
double og_sun_lon(double t, const double* par, double* fd);

/* ---- the ocean's column processes (ocean.c; reference kernels/ocean.cu) ---- */
void og_ocn_overturn(int N, int L, const double* h, double* T, double* S, double grav, double alpha, double alpha2,
                     double beta, const double* land);
void og_ocn_column(int N, int L, const double* h, double* T, double* S, const double* taum, const double* qns,
                   const double* qfw, const double* qsun, double grav, double alpha, double alpha2, double beta,
                   double S_ref, double alpha_ref, double rho_w, double cw, double m_kt, double n_kt, double hmin,
                   double hmax, double dTmin, double R, double z1, double z2, double dt, double* wx,
                   const double* land, double kappa_d, double dwf, const double* shear, double ri_b, double reach,
                   double h_ent, const double* Tc, double det_class, const double* Sc);
void og_ocn_xch_salt(int N, int L, const double* wx, const double* h, double* S, int seq);
void og_ocn_pen(int N, int L, const double* h, double* T, const double* qocean, const double* qsw, double rho_cw,
                double R, double z1, double z2, double dt, double* qns_acc, double* sun_acc, const double* land,
                const double* up);
void og_ocn_fresh(int N, const double* h, double* S, const double* evap, const double* rain, const double* ice,
                  double* ice_prev, double S_ref, double rho_w, double dt, double* fw_acc, const double* land,
                  double* mouth, double fw_local, double* hw, double* M, const double* area);
void og_ocn_diffuse(const og_grid* g, int L, const double* h, const double* T, const double* S, double* Tn,
                    double* Sn, const double* wall, double kappa_h, double dt);

/* ---- the land's water routing and ice flow (reference: kernels/land.cu, land.py:33-60); land.c ---- */
void og_land_route(const og_grid* g, const double* b, const double* land, const double* soilcap, double* pond,
                   double* soil, const double* ice, double rho_w, double rho_ice, double u, double ks, double* mouth,
                   double* qout);
void og_ice_flow(const og_grid* g, const double* b, const double* land, double* ice, double rho_ice, double gam,
                 double* ice_prev, double* calved, double* iout);

#endif
