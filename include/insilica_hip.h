/* insilica_hip.h -- the C ABI of libinsilica_hip.so: the stacked-layer TRiSK dynamical core with its split-explicit
 * RK3 step, FCT tracer transport, the one-layer SWE solver and the air-sea coupling of InSilica World V2, as
 * hand-written gfx950 HIP kernels.
 *
 * The reference drives these operations from Python through CuPy RawKernel launches (src/insilica/gpu.py:41-83); here
 * each Python driver method of the reference (Layers.step, Transport.step_many, SWE.step, Ocean.couple, ...) is one
 * entry point, which launches this library's own (fused) kernels in order on the given stream.
 *
 * Conventions (the reference's own, so tests and diagnostics read the same arrays):
 *   - layered arrays are layer-major, x[l * n + i]; transport rows are y = tracer * L + layer, (rows, N);
 *   - index tables are int32, padded with -1 to width 6 per cell and 3 per vertex; signs are int8;
 *   - every pointer is a device pointer; optional arrays are NULL, as the reference passes NULL;
 *   - nothing allocates and nothing synchronises: every buffer comes in, all work goes on `stream` (a hipStream_t;
 *     NULL is the default stream), so a caller may capture any sequence of calls in a hipGraph (the one exception is
 *     the mesh build's isw_mesh_lloyd_device, which reads its convergence back once per chunk of iterations);
 *   - every entry point returns 0, a hipError_t code (launch failure), or ISW_EBADARG for sizes or options it does not
 *     support (checked on the host before anything is launched).
 *
 * Precision: `prec` ISW_FP64 is the plain form (every operation in FP64, no contraction: the reference's "double"
 * path, --fmad=false); ISW_MIXED is the reference's mixed form (src/insilica/gpu.py:16-19,46-47,77-80): divisions by
 * fixed grid quantities become products with stored reciprocals, the PV flux gathers FP32 packed neighbour records
 * with FP32 weights, the transport stencil runs on FP32 weights and an FP32 copy of the field, and multiply-adds may
 * fuse; the state, the fast sub-step accumulation, the mass updates and the sums stay FP64.
 */
#ifndef INSILICA_HIP_H
#define INSILICA_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ISW_FP64 0
#define ISW_MIXED 1
#define ISW_EBADARG 9001
#define ISW_LMAX 8
#define ISW_KSOIL 16
#define ISW_PAIR_THREADS 256 /* threads per block of the two-sub-step patch kernel: the most edges a patch may own,
                              * and the most cells of patch and halo */

/* The mesh on the device (reference: src/insilica/gpu.py:55-80 DeviceGrid; tables from src/insilica/grid/trisk.py).
 * rA, rde, rAV: 1/areaCell, 1/de, 1/areaV; wf, d2wtf: FP32 copies of w and d2wt (read by the mixed form only).
 * d2idx/d2wt: the transport's curvature stencil, 8 per edge (reference: src/insilica/gpu.py:86-108). */
typedef struct isw_grid {
  int32_t N, E, NV, n_eoe;
  const int32_t *eI, *eJ, *nbr, *cellEdge, *cellEdgeAsc, *ev1, *ev2, *vEdge, *vCell, *eoeStart, *eoe, *d2idx;
  const int8_t *cellSign, *vEdgeSign, *vSlot;
  const double *le, *de, *areaCell, *areaV, *fV, *kite, *w, *d2wt;
  const double *rA, *rde, *rAV;
  const float *wf, *d2wtf;
  /* Patches for the air's fast sub-steps in one launch each (optional: ph_off NULL keeps two launches per sub-step;
   * no counterpart in the reference). Patch b (of npatch) holds the cells [b pcells, (b + 1) pcells); its halo cells
   * (the second cells of its edges outside it) ph_cell[ph_off[b] .. ph_off[b + 1]); its edges (those whose first cell
   * eI is in it) pe_edge[pe_off[b] .. pe_off[b + 1]), each with its cells' slots pe_slot = sI | sJ << 16 (a cell of
   * the patch at its offset in it, a halo cell at pcells + its index in the halo); pslots: the most slots a patch uses;
   * pown: the most edges a patch owns (max of pe_off[b + 1] - pe_off[b]). */
  int32_t npatch, pcells, pslots, pown;
  const int32_t *ph_off, *ph_cell, *pe_off, *pe_edge, *pe_slot;
  /* Two air fast sub-steps per launch on the same patches (optional: ph2_off NULL keeps one per launch). The first
   * sub-step's cells: patch, halo and the halo's halo ph2_cell[ph2_off[b] ..] (slots after the halo's); its winds on
   * every edge of patch and halo: the patch's edges (slots 0 .. ne - 1) and px_edge[px_off[b] ..] (slots from ne),
   * with their cells' slots px_slot (as pe_slot); pc_eslot[3 pc_off[b] ..]: each patch-then-halo cell's six edge
   * slots, two per int32 (0xffff: no edge); pslots2, pedges2: the most cell and edge slots a patch uses. The pair
   * kernel steps each owned edge on its own thread of a 256-thread block, so these tables are refused (ISW_EBADARG)
   * when a patch owns more than 256 edges (pown > 256) or its patch and halo hold more than 256 cells (pslots > 256). */
  int32_t pslots2, pedges2;
  const int32_t *ph2_off, *ph2_cell, *px_off, *px_edge, *px_slot, *pc_off, *pc_eslot;
  /* The transport's fused flux-and-stage passes on the same patches (optional: pi_off NULL keeps a flux pass and a
   * stage pass; needs pcells 64). A patch's edges are its own (pe_edge, slots 0 .. ne - 1) and those whose second
   * cell only is in it, pi_edge[pi_off[b] .. pi_off[b + 1]) (slots from ne); tc_slot (N x 6): for each cell, in
   * cellEdgeAsc order, its edge's slot in its patch times 2, plus 1 where the cell is the edge's first (-1: no edge);
   * ptedges: the most edges a patch uses (ne + its incoming edges). */
  int32_t ptedges, pad3_;
  const int32_t *pi_off, *pi_edge, *tc_slot;
  /* eslot (E x 2): each edge's slot in its first and its second cell's cellEdge ring (required with ph_off: the
   * stage's coefficients are written at those slots) */
  const int8_t *eslot;
} isw_grid;

/* A batch of worlds stepped together on one mesh: the disjoint union of the worlds' meshes (all of one level, so of
 * one step), cell c of world w at c + the world's cell offset, likewise edges, vertices and every index table. The
 * per-world values that differ between planets live in small device tables; everything else (settings, scratch) is
 * the layer stack's. Each launch then covers every world at once. Host struct, device pointers:
 *   wc (N), we (E): the world of each cell and edge of the union; grav, nu4 (nworlds): gravity and the del^4
 *   coefficient; nsub (nworlds x 3): fast sub-steps per stage; tau, inv_n (nworlds x 3): their length c dt / n and
 *   1 / n (computed as the reference's driver computes them, layers.py:276-277); nmax (3, host): the largest nsub per
 *   stage. No counterpart in the reference, whose Ensemble runs one graph per world (ensemble.py:17-23). */
typedef struct isw_batch {
  int32_t nworlds, nmax[3];
  const int32_t *wc, *we, *nsub;
  const double *grav, *nu4, *tau, *inv_n;
} isw_batch;

/* One stack of layers and its preallocated scratch (reference: src/insilica/layers.py:161-218 Layers.__init__).
 * air != 0: air layers (h is the mass per area, theta per layer); else Boussinesq layers (h in m).
 * gp (L-1), drag (L), stress (L) are values in the struct. fV (NV): the Coriolis parameter with the rotation applied.
 * b (N): the bottom. wall (E, nullable): 0 on the edges no water crosses. Gc (N): the fast mode's coefficient per cell.
 * Scratch sizes -- (L,N): B, scale, h1, h2, D1, hold, pibar, S, muf, mup, phi; (L,NV): q, Z1;
 * (L,E): Tp, F, u1, u2, L1, L2, hedge, uf, accL; (E): G, He, ubs, ub0, ub, acc, diss; (N): eta0, eta, eta_prev;
 * ufb (L,E) and ubb (E): the second buffers of the fast sub-steps' winds (the sub-steps alternate between two);
 * ubc (E, nullable): a third barotropic wind buffer, which lets the ocean's sub-steps run in one launch each;
 * pk: 3*L*E doubles (plain form); pkf: 3*L*E floats (mixed form; NULL otherwise).
 * batch: NULL for one world; else the batch the union mesh holds (then nu4 != 0 marks that some world has del^4, and
 * isw_layers_step's nfast is ignored in favour of the batch's per-world sub-steps). */
typedef struct isw_layers {
  int32_t L, air, prec, pad_;
  double grav, cp, kappa, p0, rd, nu4, floor_, div_damp, thin, shelf, vmix, vmix_cd, thin_wind;
  double gp[ISW_LMAX], drag[ISW_LMAX], stress[ISW_LMAX];
  const double *fV, *b, *wall;
  double *Gc;
  double *B, *q, *Tp, *F, *scale, *G, *He, *ubs, *ub0, *ub, *acc, *eta0, *eta, *eta_prev, *h1, *h2, *u1, *u2;
  double *D1, *Z1, *L1, *L2, *hold, *hedge, *diss, *pibar, *S, *muf, *mup, *phi, *uf, *accL, *pk, *ufb, *ubb, *ubc;
  float *pkf;
  const isw_batch* batch;
  /* Tides on Boussinesq layers (reference: layers.py:57-58, set by tides.Tides; NULL / 0 without): tide (N), the
   * tide-generating potential (m^2/s^2) subtracted from every layer's Bernoulli function; cbot (E), the internal
   * tide's drag (m/s) and cd_bot, the quadratic drag coefficient, both added to the bottom layer's stress (cd_bot
   * times |u| from the edge's normal and TRiSK tangential winds). */
  const double *tide, *cbot;
  double cd_bot;
  /* The air's fast sub-steps' coefficients per cell slot and layer, sign * l_e * h_e at [(layer * N + cell) * 6 +
   * slot] (L*N*6), written once per stage with the edge thicknesses: cfc (doubles, plain FP64 form) or cfcf (floats,
   * mixed form); the form's one is required for air layers when the grid has patch tables. */
  double *cfc;
  float *cfcf;
} isw_layers;

/* FCT transport scratch for up to rows_max rows (reference: src/insilica/transport.py:19-30 Transport.__init__).
 * (rows,N): td, p1, p2, Rp, Rm; (rows,E): Hf, A; pf: (rows,N) floats for the mixed form's stencil copy, else NULL. */
typedef struct isw_transport {
  int32_t L, rows_max, prec, pad_;
  double beta;
  double *td, *p1, *p2, *Rp, *Rm, *Hf, *A;
  float *pf;
  float *pf2; /* (rows,N) floats: the mixed form's second stencil copy, which the fused stages alternate with pf
               * (NULL: the flux and stage passes run apart) */
} isw_transport;

/* Layers.step (reference: src/insilica/layers.py:311-338; kernels/layers.cu): one time-split RK3 step in place on
 * h (L,N) and u (L,E); theta (L,N) for air, else NULL; Fm (L,E, nullable) receives the last stage's l*dt*F.
 * Stage s takes ceil(c_s * nfast) fast sub-steps (c = 1/3, 1/2, 1). Then the interfacial stress (vmix). */
int32_t isw_layers_step(const isw_grid* g, const isw_layers* lay, double* h, double* u, double* theta, double* Fm,
                        double dt, int32_t nfast, void* stream);

/* Layers._lap (reference: layers.py:249-253; kernels/trisk.cu:14-45): out = grad(div u) - k x grad(curl u), (L,E). */
int32_t isw_layers_lap(const isw_grid* g, const isw_layers* lay, const double* u, double* out, void* stream);

/* Layers.update_gfast (reference: layers.py:230-235; kernels/layers.cu:234-250 air_gfast): lay->Gc from h, theta. */
int32_t isw_layers_gfast(const isw_grid* g, const isw_layers* lay, const double* h, const double* theta, void* stream);

/* Layers.exchange (reference: layers.py:340-358; kernels/layers.cu:454-565,599-621): move wx (L-1,N) across the
 * interfaces with tracer psi (L,N), volume M (L,N) and momentum; theta (nullable) for air, which takes the kinetic
 * energy lost as heat along with nu4dt times the last step's del^4 loss; seq != 0 takes donors in passing order. */
int32_t isw_layers_exchange(const isw_grid* g, const isw_layers* lay, const double* wx, double* h, double* u,
                            double* M, double* psi, double* theta, int32_t seq, double nu4dt, void* stream);

/* vmix_heat (reference: kernels/layers.cu:599-621; world.py:317-318): each edge's kinetic-energy loss diss (E) back
 * as heat, half to each cell, spread over the air column by mass. */
int32_t isw_vmix_heat(const isw_grid* g, const isw_layers* lay, const double* h, const double* diss, double* theta,
                      void* stream);

/* Transport.mass_after (reference: transport.py:32-35; kernels/trisk.cu:48-59): M1 = M0 moved by Fm, (L,N). */
int32_t isw_mass_update(const isw_grid* g, int32_t L, const double* M0, const double* Fm, double* M1, void* stream);

/* Transport.step_many (reference: transport.py:40-71; kernels/fused.cu:124-241): the rows tracer rows stacked in psi
 * (rows, N; rows a multiple of tr->L, row y moved by layer y % L's fluxes) moved in place by the mass fluxes Fm (L,E)
 * from M0 to M1 (L,N): 3rd-order fluxes in RK3 stages with Zalesak FCT (order 3), or upwind (order 1). */
int32_t isw_transport_step(const isw_grid* g, const isw_transport* tr, const double* M0, const double* M1,
                           const double* Fm, double* psi, int32_t rows, int32_t order, void* stream);

/* SWE.step (reference: src/insilica/swe.py:124-132; kernels/fused.cu:27-120): Wicker-Skamarock RK3 with the
 * positivity limiter, L independent layers. Scratch from `lay` (F, B, q, scale, h1, h2, u1, u2, D1, Z1, L1, L2);
 * grav, drag (L), nu4, b (N) from `lay`; Fm (L,E, nullable) receives the last stage's l*dt*F. */
int32_t isw_swe_step(const isw_grid* g, const isw_layers* lay, double* h, double* u, double* Fm, double dt,
                     void* stream);

/* Ocean.couple (reference: src/insilica/ocean.py:193-202; kernels/ocean.cu:17-42 couple_stress): the exact wind
 * stress between the air's bottom layer (La layers: ha (La,N), ua (La,E)) and the ocean's top layer (ho, uo);
 * diss (E) gets the kinetic energy removed, tau_acc (E) += tau^2 dt; wall (E) nullable; cd (N) nullable. */
int32_t isw_couple_stress(const isw_grid* g, int32_t La, const double* ha, double* ua, const double* ho, double* uo,
                          double c, double rho_w, double dt, double* diss, double* tau_acc, const double* wall,
                          double land_drag, const double* cd, void* stream);

/* Ocean.ground's land column (reference: ocean.py:177-182; kernels/ocean.cu:434-465 land_column): implicit soil
 * conduction under the skin T[0] (N) of land cells, K <= 16 layers Tg (K,N) of thickness dz (K, device), one FP64
 * Thomas solve per column (both forms). */
int32_t isw_land_column(const isw_grid* g, int32_t K, const double* land, double* T, const double* cap, double* Tg,
                        const double* dz, double c_soil, double k_soil, double dt, void* stream);

/* Ocean.update_cap (reference: ocean.py:188-191; kernels/ocean.cu:468-471 ocn_cap). land nullable. */
int32_t isw_ocn_cap(const isw_grid* g, const double* h, double rho_cw, double* cap, const double* land, double c_land,
                    void* stream);

/* gpu.dsum (reference: src/insilica/gpu.py:111-126; kernels/trisk.cu:61-81 block_sum): the fixed-order tree sum of
 * x (or of x*y, y nullable) over n values into out[0]: 256 threads x 8 values per block, a halving tree, passes
 * repeated to one block. scratch: at least ceil(n / 2048) + ceil(n / 2048^2) + 2 doubles. */
int32_t isw_dsum(const double* x, const double* y, int64_t n, double* scratch, double* out, void* stream);

/* isw_dsum without a workspace argument. No counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc). It launches
 * isw_dsum's kernels in the same fixed-order tree, so out[0] is bit-identical to isw_dsum's, with the workspace in a
 * file-scope device array of ISW_DSUM_NOALLOC_DOUBLES doubles (131 KB, held on each device the library runs on
 * until the process ends); hence n <= ISW_DSUM_NOALLOC_MAX_N (else ISW_EBADARG). It calls no allocator. The first
 * call on a device looks the array's address up there (hipGetSymbolAddress, a host query). Every call on a device
 * shares that device's array, which keeps the last call's partial sums: calls that overlap on one device (two
 * streams, two host threads) race, and the caller must order them. Nothing in the array needs to start at zero:
 * each pass writes its partials before the next reads them. */
#define ISW_DSUM_NOALLOC_MAX_N 33554432 /* 2^25 values */
#define ISW_DSUM_NOALLOC_DOUBLES 16400
int32_t isw_dsum_noalloc(const double* x, const double* y, int64_t n, double* out, void* stream);

/* A compensated, deterministic sum of x (or x*y) into out[0]: per-block Neumaier partial sums with their error terms
 * in a fixed order, then one fixed-order compensated pass over the partials; out[0] the sum, out[1] its remaining
 * correction (out holds 2 doubles). scratch: 2 * ceil(n / 2048) doubles. The budgets' sum (no counterpart in the
 * reference, whose totals use dsum). */
int32_t isw_csum(const double* x, const double* y, int64_t n, double* scratch, double* out, void* stream);

/* ---- the ocean's column processes (reference: src/insilica/kernels/ocean.cu; ocean.py:226-295; FP64 in both forms) ----
 *
 * The ocean's settings these read (reference: ocean.py:29-33 OCEAN, an Ocean's cfg; ALPHA_REF ocean.py:27; the
 * equation of state b = grav (alpha th + alpha2 th^2 / 2 - beta S), th = T - 273.15 K, ocean.cu:43-51). Host struct. */
typedef struct isw_ocean {
  double grav, alpha, alpha2, beta, S_ref, alpha_ref, rho_w, cw, m_kt, n_kt, hmin, hmax, dTmin, sw_R, sw_z1, sw_z2,
      kappa_d, dwf, ri_b, reach, h_ent, det_class, fw_local;
} isw_ocean;

/* ocn_overturn (reference: ocean.cu:231-237, adjust 55-72): convective adjustment of every sea column's T and S (L, N)
 * on thicknesses h (L, N): a layer denser than the one below mixes with it completely, repeated until the column is
 * stable; land (N, nullable) cells > 0.5 are skipped. 2 <= L <= ISW_LMAX. */
int32_t isw_ocn_overturn(int32_t N, int32_t L, const isw_ocean* o, const double* h, double* T, double* S,
                         const double* land, void* stream);

/* The mixing interval's forcing (reference: Ocean.step, ocean.py:257-266): tau_acc (E; the sum of tau^2 dt) becomes
 * the edges' rms stress sqrt(tau_acc (1 / dt)) in place and taum (N) its cell value (air.cu:1978-1989's cell speed);
 * q_acc, sun_acc, fw_acc (N) are divided by dt in place. */
int32_t isw_ocn_forcing(const isw_grid* g, double dt, double* tau_acc, double* taum, double* q_acc, double* sun_acc,
                        double* fw_acc, void* stream);

/* ocn_column (reference: ocean.cu:117-213): the mixed layer over one mixing interval dt: adjustment, Kraus-Turner
 * entrainment or shoaling on the interval's rms stress taum and mean fluxes qns (non-solar heat), qfw (fresh water),
 * qsun (sunlight) (N); shear (N, nullable) the base's velocity jump for ri_b; the interior interfaces' deep-water
 * formation (dwf) and mixing (kappa_d); reach, h_ent; Tc, Sc (L, nullable) the layers' classes for det_class. Writes
 * T, S (adjusted) and wx (L - 1, N), the water to move up across each interface (for isw_layers_exchange). */
int32_t isw_ocn_column(int32_t N, int32_t L, const isw_ocean* o, const double* h, double* T, double* S,
                       const double* taum, const double* qns, const double* qfw, const double* qsun, double dt,
                       double* wx, const double* land, const double* shear, const double* Tc, const double* Sc,
                       void* stream);

/* ocn_xch_salt (reference: ocean.cu:352-381): S (L, N) moved with the water wx (L - 1, N) moves, before the exchange
 * moves the rest; upwind; seq != 0: in passing order. */
int32_t isw_ocn_xch_salt(int32_t N, int32_t L, const double* wx, const double* h, double* S, int32_t seq, void* stream);

/* ocn_pen (reference: ocean.cu:387-407; Ocean.accumulate, ocean.py:214-224): the sunlight qsw (N, W/m^2) passing the
 * mixed layer's base heats the layers below over dt; qns_acc += (qocean - qsw) dt, sun_acc += qsw dt; up (N,
 * nullable): where positive the mixed layer keeps all the sunlight. */
int32_t isw_ocn_pen(int32_t N, int32_t L, const isw_ocean* o, const double* h, double* T, const double* qocean,
                    const double* qsw, double dt, double* qns_acc, double* sun_acc, const double* land,
                    const double* up, void* stream);

/* ocn_fresh (reference: ocean.cu:245-261): rain - evap (kg/m^2/s) and the sea ice's growth (ice - ice_prev, kg/m^2)
 * freshen or salt the mixed layer's S[0] over dt (at S_ref, or with fw_local > 0 with its volume: h_w, M (N) of the
 * mixed layer rewritten from area (N)); mouth (N, nullable) emptied into it; ice_prev = ice; fw_acc += the water. */
int32_t isw_ocn_fresh(int32_t N, const isw_ocean* o, const double* h, double* S, const double* evap,
                      const double* rain, const double* ice, double* ice_prev, double dt, double* fw_acc,
                      const double* land, double* mouth, double* hw, double* M, const double* area, void* stream);

/* ocn_diffuse (reference: ocean.cu:268-292): T, S (L, N) mixed along each layer over dt into Tn, Sn (L, N); wall (E,
 * nullable) closes edges. */
int32_t isw_ocn_diffuse(const isw_grid* g, int32_t L, const double* h, const double* T, const double* S, double* Tn,
                        double* Sn, const double* wall, double kappa_h, double dt, void* stream);

/* Fill n doubles with v (a kernel, not hipMemsetAsync, so that captured graphs replay it safely). */
int32_t isw_fill(double* x, int64_t n, double v, void* stream);

/* ---- k-table radiation (reference: src/insilica/radiation.py, kernels/air.cu:2111-2511) ----
 *
 * The column's parameters arrive as the reference's `par` array (atmosphere.PAR_ORDER, 132 values, then the theta*
 * of the interfaces), in HOST memory: the entry points read eps_s [4], grav [13], cp [14], kappa [15], p0 [16],
 * tilt [17], year_s [18], day_s [19], cloud_g [59], bl_q [89], edge_k [101], ice_air [102], lw_refl [124], ecc [126]
 * and lam_peri [127] and pass them to the kernels by value. Layers k = 0 (top) .. L-1 (bottom), L <= ISW_LMAX; each
 * is split into 4 sub-levels, s = 4k + j. The sub-levels' properties live in `sub`, float32, sub[(s*15 + v)*N + i]
 * (v: 0/1 edge temperatures top/bottom, 2/3 table positions in log p and T, 4-8 H2O/CO2/O3/N2O/CH4 molecules per
 * cm^2, 9 H2O's place between its two tables, 10 the layer's cloud longwave optical depth in its cloud sub-level,
 * 11 cloud cover, 12 air molecules per cm^2, 13 the haze's share, 14 collision-induced absorption's weight). */

/* The air's radiative constants (reference: the AIR_* compile-time defines of kernels/air.cu:75-116 that
 * constants.air_props sets for another air; Earth air: m_air 28.9647, m_air_kg 28.9647e-3, o2 0.2095, ray = broad =
 * o3 = 1, eps 0.622 (water's molar mass over the air's, in the saturation humidity), rd 287.0 (the column physics' gas
 * constant), vf 0.61 (the virtual-temperature factor), cia 0). cia != 0: lw_setup writes the collision-induced
 * absorption weights (v = 14). */
typedef struct isw_air {
  double m_air, m_air_kg, o2, ray, broad, o3, eps, rd, vf;
  int32_t cia, pad_;
} isw_air;

/* Longwave tables (reference: Longwave.__init__, radiation.py:26-45): logk (NC = 6, NP, NT, NG) ln k in
 * cm^2/molecule, component-major; planck (NTP, NG) per-g Planck radiance; o3 (NO3) ozone mole fraction on the
 * pressure axis; lp0 = ln P[0], dlp the log-pressure step, t0, dtt the temperature axis, tp0 the Planck table's first
 * temperature, o3lp0/o3dlp the ozone axis; cia (NT, NG) or NULL. Device pointers. NG <= 1024. */
typedef struct isw_lw_tables {
  int32_t NC, NP, NT, NG, NTP, NO3;
  double lp0, dlp, t0, dtt, tp0, o3lp0, o3dlp;
  const float *logk, *planck, *o3, *cia;
} isw_lw_tables;

/* Shortwave tables (reference: Shortwave.__init__, radiation.py:89-106): logk (NC = 7, NP, NT, NG); sun (NG) the
 * solar weights summing to 1; ray (NG) Rayleigh cross sections per air molecule; hzf (NG) the haze's spectral factor;
 * cssa (NG) the cloud drops' single-scattering albedo; cssa_i, g_i (NG) ice clouds' optics or NULL; cia (NT, NG) or
 * NULL; yscale maps lw_setup's temperature position to this table's; tau550, hssa, hg the haze's optical depth, albedo
 * and asymmetry; inhom the cloud inhomogeneity factor. Device pointers. NG <= 1024. */
typedef struct isw_sw_tables {
  int32_t NC, NP, NT, NG;
  double yscale, tau550, hssa, hg, inhom;
  const float *logk, *sun, *ray, *hzf, *cssa, *cssa_i, *g_i, *cia;
} isw_sw_tables;

/* Longwave.__call__'s first kernel (reference: kernels/air.cu:2145-2229 lw_setup): the sub-levels of every
 * (column, layer) from the layer masses mu (L, N), theta (L, N), vapor qv (L, N; NULL dry), the clouds' cloudo
 * ((4L + 3), N; NULL clear: row k cover, row L + 3 + k in-cloud longwave optical depth), the gas tracers gx
 * (3, L, N; NULL: the planet's co2, n2o, ch4), the haze scale height hs (m) and the temperature mode tmode (0
 * isothermal, 1 uniform theta, 2 standard lapse) -> sub. prec: ISW_FP64 or ISW_MIXED. */
int32_t isw_lw_setup(int32_t N, int32_t L, const double* par, const isw_lw_tables* t, const isw_air* air, double co2,
                     double n2o, double ch4, const double* gx, const double* mu, const double* theta, const double* qv,
                     const double* cloudo, double hs, int32_t tmode, float* sub, int32_t prec, void* stream);

/* Longwave.__call__'s second kernel (reference: kernels/air.cu:2231-2348 lw_column): per column, per g-point, the
 * optical depths from the tables, two-stream down and up through the 4L sub-levels (clear and cloudy streams mixed per
 * layer), the surface at Ts (N). Writes radf ((3(L+1) + 3 + 6L), N): rows 0..L up (air's emission) and L+1..2L+1
 * down at the layer edges, 2(L+1)..3(L+1)-1 the surface emission's share reaching each edge, 3(L+1) the clear sky's
 * top up flux, +1 its transmission share, +2 the surface's Planck sum, then 6 rows per layer at its edge sub-levels.
 * Sums over g-points in FP64 of per-g values rounded to float32, as the reference's. */
int32_t isw_lw_column(int32_t N, int32_t L, const double* par, const isw_lw_tables* t, const float* sub,
                      const double* Ts, double* radf, int32_t prec, void* stream);

/* Shortwave.__call__ (reference: kernels/air.cu:2350-2511 sw_column, sw_layer): per column, per g-point, gases,
 * Rayleigh, haze and clouds per sub-level, Delta-Eddington, adding over the surface albedo alb (N). The sun: mode 0 the
 * mean of its cosine over the interval's daylight from the device clock t (1 double, s) and the unit cell positions
 * pos (N, 3); mode 1 a fixed 0.5. Writes swf ((3L + 2), N) per unit of sunlight on the horizontal: rows 0..L-1 each
 * layer's absorbed share, L the surface's, L + 1 the clear sky's reflected share, then each layer's top and bottom
 * sub-level's absorption. cloudo as lw_setup's (rows k cover, 2L + 3 + k shortwave optical depth, 3L + 3 + k the
 * liquid share) or NULL. */
int32_t isw_sw_column(int32_t N, int32_t L, const double* par, const isw_sw_tables* t, const isw_air* air,
                      const float* sub, int32_t mode, const double* tnow, double interval, const double* pos,
                      const double* alb, const double* cloudo, double* swf, int32_t prec, void* stream);

/* ---- the air column's Held-Suarez path and the surface exchange (reference: kernels/air.cu; FP64 in both forms) ---- */

/* AirPhysics.column with the default forcing (reference: atmosphere.py:134-158; kernels/air.cu:876-1972 air_column):
 * the sun (mode 0 the diurnal and seasonal sun from the device clock t (1 double, s) at the unit cell positions pos
 * (N, 3); mode 1 the mean insolation insol (N)); the shortwave absorbed in the air and at the surface (the k-tables'
 * swf ((3L + 2), N) when given, else the gray split); the gray longwave (or the k-tables' radf ((9L + 6), N)); the
 * slab surface at Ts (N; heat capacity cap (N) when given, else c_slab) under the bulk exchange or the surface
 * layer's rho C_H U (ch (N), with sfc > 0); the heating as theta and the crossing of the interfaces (the k-tables'
 * edge heating). With moist > 0 also the vapor qv (L, N) and sea ice ice (N, kg/m^2): evaporation, the ice's skin
 * and albedo, freezing and melting, condensation per layer, moist convection (the parcel plume; with plume_deep > 0
 * the deep plume's full ascent, mixing (plume_ent) and entraining the cloud layer's air (ent_cloud, ent_det), its
 * detrained condensate as cloud (l_det), frozen share (ice_air), lightning into flash and downdrafts (downdraft)),
 * and the vapor's longwave part with lw_w0. With clouds (moist, cloud > 0 and the cloud water cloudw (L, N) given,
 * updated in place): the cloud statistics and erosion, Sundqvist's autoconversion, the ice's fall (ice_air: its
 * freezing, melting and snow), the rain's evaporation (k_revap, or Kessler's with revap; conv_revap below the plumes),
 * the clouds' reflection and emission, the cloud water carried by the crossing; cloudo (L + 3 rows, or 4L + 3 with
 * radf, N) when given: each layer's cover, the total cover, the clouds' shortwave and longwave effects at the top
 * (and with radf the in-cloud longwave and shortwave optical depths and liquid share), as the radiation's set-up reads
 * them. Over land (moist; land (N) > 0.5 marks a land cell; soilcap (N) the soil bucket's capacity, kg/m^2): the
 * soil water soil, the ponds pond and the snow snow (N, kg/m^2, updated in place), the ground's albedo by its
 * wetness, the snow's albedo, the surface resistance (land_rs, rs_max, bare_fx), the ground's implicit skin (c_deep),
 * the ground's capacity by its water (c_wet: cap rewritten), land_exch, firn on the land's ice sheets (firn), rain
 * and snow landing, ponds freezing and soaking in (k_inf), a full bucket springing; lc (8, N) the land cover's rows
 * when given (row 0 the roughness, read by isw_surface_layer). With bl_mix > 0 the bottom layer is the mixed boundary
 * layer: the surface's fluxes go into it alone, the lowest interface floats with it, it entrains the air above by
 * the surface's buoyancy flux (bl_ent; bl_reach: dry convection into the layers above colder than it) and gives its
 * top back at night down to mu_night, ends at its condensation level with bl_lcl (ent_lcl: its entrainment not
 * stopped there), and a layer emptied to its floor keeps its theta within its edges. The moist options sat_prof (a
 * middle layer's capacity with its theta spread between its edges), sat_base (the third layer condensing at its
 * base), edge_vap (air crossing an edge carries the vapor there), freeze_dry (air rising into the stratosphere holds
 * at most the cold point's saturation), q_shape (the surface air's vapor at the bottom layer's relative humidity) and
 * plume_root with root_q (plumes rooted in the surface air, with the thermals' excess) as the reference's. Updates
 * theta, Ts (and qv, ice) in place and writes wx (L-1, N), olr, sw_abs, stored (the column's energy change, J/m^2),
 * rain, evap, conv
 * (kg/m^2/s), tsurf; rainp (4, N: rain from the layers above the bottom one, from the bottom one, the plumes', snow),
 * qocean, qsw, flash, swalb and firn (the firn's temperature) when given; cap (N) the surface's heat capacity, else
 * c_slab. The argument list is the reference's; elev (N) with theta_rise > 0 raises the lowest interface over high
 * ground. dgo (5 L - 2, N, nullable): the diagnostics rows (air.cu:1286-1290, 1944-1950): each layer's heating Q and
 * latent heating (W/m^2), each interface's gross up and down dry crossing (kg/m^2 per step, before netting and
 * limiting), each layer's convective mass delivered into it. Returns ISW_EBADARG for moist air without qv or ice, or
 * with land but without soilcap, soil, pond or snow. bl_q (the mixed layer's even vapor and its
 * cloud statistics; ml_clear, cu_cloud), the deep plume's IFS entrainment and detrainment (ent_ifs, ent_loc, det_all),
 * virtual buoyancy (buoy_v), the shallow plume (shallow, rooted by plume_root with root_q) and the cumulus layer
 * (cu_dth's floating edges with cu_dp and cl_float, the plumes' air taken in through it moved with them, cu_sub's
 * subsidence with sub_mix; air.cu:668-713, 913-922, 1601-1700, 1748-1861) as the reference's. xfer (3 L, N,
 * nullable): the gas tracers' mole fractions per layer, carried by every move of the column's air (the reference's
 * -DGASX column, air.cu:503-507), updated in place.
 * prec: ISW_FP64, or ISW_MIXED (the reference's mixed build: 1 - e^-x, the edge sub-levels' Exner powers and the
 * saturation and condensation arithmetic in float32). */
int32_t isw_air_column(int32_t N, int32_t L, const double* par, const isw_air* air, int32_t mode, const double* t,
                       const double* insol, const double* pos, const double* uabs, const double* mu, double* theta,
                       double* qv, double* Ts, double* ice, double dt, double* wx, double* olr, double* sw_abs,
                       double* stored, double* rain, double* evap, double* conv, double* tsurf, double* rainp,
                       double* dgo, double* cap, double* qocean, double* qsw, const double* land,
                       const double* soilcap, double* soil, double* pond, double* snow, double* cloudw, double* cloudo,
                       const double* elev, double* flash, const double* radf, const double* swf, double* swalb,
                       const double* ch, const double* lc, double* firn, double* xfer, int32_t prec, void* stream);

/* AirPhysics.column with forcing='held_suarez' (reference: atmosphere.py:142-146; kernels/air.cu:2065-2109 hs_column
 * with the interface crossing 744-850): each of L >= 2 layers relaxed toward Held & Suarez's equilibrium theta at its
 * mass-mean pressure (rates hs[0] = k_a, hs[1] = k_s, 1/s; HOST array of 2), lat (N) the cells' latitudes; then the
 * air crossing each isentropic interface theta* when heated or cooled across it. theta (L, N) updated in place;
 * wx (L-1, N) the mass to move up across each interface (for isw_layers_exchange); qnet (N) the heating (W/m^2);
 * stored (N) the column's enthalpy change (J/m^2). mu (L, N) is read only. par as the radiation's (host), with the
 * interfaces' theta* at par[132..]; bl_mix > 0 (the mixed bottom layer) needs L >= 3. */
int32_t isw_hs_column(int32_t N, int32_t L, const double* par, const double* hs, const double* lat, const double* mu,
                      double* theta, double dt, double* wx, double* qnet, double* stored, void* stream);

/* AirPhysics.surface (reference: atmosphere.py:125-132; air.cu:2015-2061 surface_layer): the surface layer's exchange
 * coefficients cd, ch (N) = rho C_D U and rho C_H U (kg/m^2/s) by Louis's bulk form from the bottom layer's wind speed
 * uabs (N), mu, theta (L, N), vapor qv (L, N; NULL dry), the surface temperature Ts (N), sea ice (N; NULL none), the
 * land mask (N; NULL all sea) and the land cover's roughness lc (N; NULL: par's z0_land). air: rd, vf, eps. */
int32_t isw_surface_layer(int32_t N, int32_t L, const double* par, const isw_air* air, const double* uabs,
                          const double* mu, const double* theta, const double* qv, const double* Ts, const double* ice,
                          const double* land, double* cd, double* ch, const double* lc, void* stream);

/* AirPhysics._speed (reference: atmosphere.py:109-115): the bottom layer's wind speed at cells from its edge winds
 * u (E): cell_speed from the edges' kinetic energy (air.cu:1978-1989), cell_speed_perot from Perot's reconstructed
 * wind with op (N, 6, 2) east/north coefficients per edge slot (air.cu:1994-2006; op from orodrag.operators). */
int32_t isw_cell_speed(const isw_grid* g, const double* u, double* out, void* stream);
int32_t isw_cell_speed_perot(const isw_grid* g, const double* u, const double* op, double* out, void* stream);

/* World._tick's gas sources (reference: world.py:343-346; air.cu:2133-2137 gas_src): x (G*L, N) gas mole fractions,
 * each gas's surface source src (G, N) (mole fraction x kg/m^2/s) into the bottom layer of mass h (L, N) over dt. */
int32_t isw_gas_src(int32_t N, int32_t L, int32_t G, const double* src, const double* h, double dt, double* x,
                    void* stream);

/* The device clock (reference: air.cu:1975 advance_time): t[0] += dt on the stream (a kernel, so captured ticks
 * advance it on replay). */
int32_t isw_advance_time(double* t, double dt, void* stream);

/* Measurement mode (no counterpart in the reference, whose tools/profile.py wraps its kernels in CUDA event pairs
 * from Python; this is that instrument inside the library). isw_timing_start() switches it on: from then every kernel
 * launched directly (not under stream capture) by the entry points above is bracketed by two HIP events on its stream
 * and tagged with its name and algorithmic bytes (each distinct array the kernel touches counted once at its stored
 * width, twice if read and written, index tables included). isw_timing_stop() switches it off, waits for the events
 * and writes one line per kernel, "name launches total_ms total_bytes", into out (cap bytes); it returns the text's
 * length, or minus the size needed if cap is too small (the records are then dropped). Host-side bookkeeping only:
 * it creates events, not device memory. Not thread-safe; not for use while a graph is being captured. */
int32_t isw_timing_start(void);
int32_t isw_timing_stop(char* out, int32_t cap);

/* Water over and through the land (reference: LandWater.step, land.py:33-46; kernels/land.cu:38-101): pond (kg/m^2)
 * runs toward lower water surfaces z = b + pond / rho_w + ice / rho_ice at 3/4 of the speed U over a depth
 * min(head, pond / rho_w), at most 1/12 of the head per edge; soil water seeps between water tables
 * zt = soil / soilcap b by Darcy's law (kseep); a donor gives at most half its pond and all its soil in a step; a sea
 * cell (land <= 0.5) receives into mouth (kg/m^2, accumulated). b, land, soilcap, ice (N); pond, soil (N) updated in
 * place; u_dt = U dt (m), kseep_dt = K dt (m^2); work: 4 N doubles; qout (N, nullable): each land cell's surface
 * outflow, accumulated (kg/m^2). */
int32_t isw_land_route(const isw_grid* g, const double* b, const double* land, const double* soilcap, double* pond,
                       double* soil, const double* ice, double rho_w, double rho_ice, double u_dt, double kseep_dt,
                       double* work, double* mouth, double* qout, void* stream);

/* The ice sheets' flow (reference: LandWater.ice_step, land.py:48-60; kernels/land.cu:104-173): Glen's shallow-ice law
 * without sliding, through each edge gam H^5 |grad s|^2 (head / de) le per step with s = b + ice / rho_ice (0 at sea)
 * and |grad s|^2 the TRiSK kinetic-energy rule's, the two ends' mean between land cells; at most 1/12 of the head per
 * edge and all the donor's ice. ice (N, kg/m^2) updated in place; gam_dt = 2 A (rho_ice g)^3 / 5 dt; work: 3 N doubles;
 * a sea cell's gain is added to calved (N) and to ice_prev (N, nullable: the ocean's sea-ice record); iout (N,
 * nullable): each land cell's outflow, accumulated. */
int32_t isw_ice_flow(const isw_grid* g, const double* b, const double* land, double* ice, double rho_ice,
                     double gam_dt, double* work, double* ice_prev, double* calved, double* iout, void* stream);

/* Lloyd's iteration toward the spherical centroidal Voronoi tessellation on the device (reference: grid/icosa.py:194-224
 * lloyd_scvt, which grid/__init__.py:95-99 runs on the GPU from level 6): pos (N, 3, device) the unit generators,
 * updated in place; ring (N, 6) int32 each cell's neighbours counter-clockwise (entries below deg[i] in [0, N)), deg (N)
 * uint8 (5 or 6); generators 0 and 11 stay. Iterates until the largest move (radians) < tol or max_iters; *iters, *moved
 * (host) the iterations done and the last largest move. work (device): at least 3 N + chunk doubles (the second iterate
 * and the chunk's moves). Launches chunk iterations at a time and reads their moves back (so it synchronises the
 * stream once per chunk); 1 <= chunk <= ISW_LLOYD_MAX_CHUNK. The fan triangles' angles use a correctly rounded
 * small-argument arctangent, so the result agrees with the host builder's to rounding. */
#define ISW_LLOYD_MAX_CHUNK 4096
int32_t isw_mesh_lloyd_device(int64_t N, double* pos, const int32_t* ring, const uint8_t* deg, int32_t max_iters, double tol,
                       double* work, int32_t chunk, int32_t* iters, double* moved, void* stream);

/* Library version and the gfx target it was built for. */
const char* isw_version(void);

/* Host-only argument checks (no counterpart in the reference, whose kernels check nothing): each returns 0 where the
 * entry points above would accept the argument and ISW_EBADARG where they refuse it, and launches nothing.
 * isw_grid_check: the mesh struct every grid-taking entry point checks first (tables present, patch tables
 * consistent, at most ISW_PAIR_THREADS owned edges and patch-and-halo cells per patch for the two-sub-step tables). isw_air_column_options:
 * the column options (par in PAR_ORDER, host) that isw_air_column does not carry for L layers. */
int32_t isw_grid_check(const isw_grid* g);
int32_t isw_air_column_options(int32_t L, const double* par);

/* The mountains' drag on the air (reference: OroDrag.apply, orodrag.py:62-68; kernels/drag.cu:40-193, Lott & Miller
 * 1997 on the isentropic layers): per cell, from its sub-grid orography sso (4, N: the heights' standard deviation mu,
 * the slope sigma, the anisotropy gamma and the axis's angle from east), the low-level flow between mu and 2 mu, the
 * blocked layer below Z_blk (its form drag, implicit: beta (L, N) = dt x rate) and the gravity waves launched at
 * 2 mu and deposited going up (Lindzen's saturation, critical levels; F (2L, N): east and north force per area per
 * layer); then per edge and layer u <- (u + dt fn / m) / (1 + beta_e), fn the cells' forces along the edge's normal
 * (nrm (E, 4): the normal's east and north at its first and second cell), m the layer's mass at the edge, the
 * kinetic energy removed added to diss (E, J/m^2). mu, theta (L, N) the air layers; u (L, E) updated in place; op
 * (N, 6, 2) Perot's coefficients for a cell's wind (east, north) per edge slot; beta, F: outputs (work). par (host):
 * grav, p0, kappa, rd, G, cd, hnc, ric, dt (equal to dt), mu_min (the reference's par vector, orodrag.py:53). L >= 2. */
typedef struct isw_oro {
  double grav, p0, kappa, rd, G, cd, hnc, ric, dt, mu_min;
} isw_oro;
int32_t isw_oro_drag(const isw_grid* g, int32_t L, const isw_oro* par, const double* mu, const double* theta,
                     double* u, const double* op, const double* nrm, const double* sso, double* beta, double* F,
                     double dt, double* diss, void* stream);

/* The gravity waves not tied to the ground (reference: WaveDrag.apply, orodrag.py:86-96; kernels/drag.cu:205-293,
 * Alexander & Dunkerton 1999): per cell a spectrum of waves, phase speeds -cmax .. cmax by dc (at most
 * ISW_GW_MAX_SPEEDS), launched at the pressure psrc centred on the source layer's wind, each deposited where it
 * breaks or meets its critical level, a layer taking no more than brings its wind to the wave's speed in a step and
 * the source layer keeping what none could take (each column's momentum exact); F (2L, N) east rows the force per
 * area, north rows 0; then the edges as isw_oro_drag's with beta = zero ((L, N) zeros). conv (N, nullable): the
 * column's convective rain for the convective source share fc; cmean (1 double, device) receives its area mean
 * (required with conv). par (host): grav, p0, kappa, rd, bt, cw, kw = 2 pi / lambda, psrc, cmax, dc, dt (equal to
 * dt), fc (orodrag.py:76-77). L >= 2. */
#define ISW_GW_MAX_SPEEDS 4096
typedef struct isw_gw {
  double grav, p0, kappa, rd, bt, cw, kw, psrc, cmax, dc, dt, fc;
} isw_gw;
int32_t isw_wave_drag(const isw_grid* g, int32_t L, const isw_gw* par, const double* mu, const double* theta,
                      double* u, const double* op, const double* nrm, const double* conv, double* cmean,
                      const double* zero, double* F, double dt, double* diss, void* stream);

/* The Sun's and the Moon's places for the tides (reference: tides.sky_par, tides.py:42-51, the same order): the
 * planet's tilt, year and day (s), orbital eccentricity and longitude of perihelion (the air's own orbit), the Sun's
 * GM and semi-major axis; the Moon's GM, semi-major axis, eccentricity, inclination to the ecliptic, sidereal month,
 * node regression and perigee advance periods (s), its longitude, node and perigee at time 0 (radians); the Love
 * factor 1 + k2 - h2; the planet's radius (m). A Moon taken away is GM 0. */
typedef struct isw_sky {
  double tilt, year_s, day_s, ecc, lam_peri, sun_gm, sun_a, moon_gm, moon_a, moon_e, moon_i, moon_T, node_T, aps_T,
      moon_L0, node0, aps0, love, radius;
} isw_sky;

/* The tide-generating potential (reference: Tides.force, tides.py:70-73; kernels/tides.cu:58-71): at each cell
 * (pos (N, 3) unit vectors) V = love sum_b GM_b R^2 / r_b^3 (3 cos^2 psi_b - 1) / 2 over the Sun and the Moon at
 * time[0] - back (time: 1 double on the device, the world's clock), into V (N). */
int32_t isw_tide_pot(const isw_grid* g, const isw_sky* tp, const double* time, double back, const double* pos,
                     double* V, void* stream);

/* The internal tide's drag (reference: Tides.force, tides.py:74-76; kernels/tides.cu:128-144): per edge
 * cbot = mean over its two cells of kappa h^2 N / 2 (rough (N): the floor's roughness h^2, m^2; N the buoyancy
 * frequency above the floor, nbot (N, nullable) or from the two lowest layers of h, T, S (L, N) with o's equation of
 * state). */
int32_t isw_tide_itd(const isw_grid* g, int32_t L, const isw_ocean* o, const double* h, const double* T,
                     const double* S, const double* rough, double kappa, const double* nbot, double* cbot,
                     void* stream);

/* Tidal mixing after an ocean step (reference: Tides.mix, tides.py:78-87; kernels/tides.cu:85-122): each sea cell's
 * bottom dissipation D = rho (c_bot + cd |u_b| + mean cbot) |u_b|^2 (u (L, E) the bottom layer's winds), D dt added to
 * dsum (N); gamma D dt exp(-z / zeta) mixes T and S (L, N) across each interface z above the floor in equal volumes
 * (heat and salt kept exactly); land (N, nullable) cells skipped. */
int32_t isw_tide_mix(const isw_grid* g, int32_t L, const isw_ocean* o, const double* h, const double* u, double* T,
                     double* S, const double* land, double c_bot, double rho, double gamma, double dt, double* dsum,
                     double cd, const double* cbot, double zeta, void* stream);

#ifdef __cplusplus
}
#endif

#endif
