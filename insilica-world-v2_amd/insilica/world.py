"""A world: a stack of layers on the layered core carrying its tracers, with an ocean under air worlds; the
reference's `World` (src/insilica/world.py:55-611) for the dynamical core, FCT transport and the air-sea coupling.

State (float64, layer-major, resident on the GPU): h (L, N), u (L, E), M (L, N) (the layer's volume or mass per cell,
moved by the same fluxes as h so tracer budgets close exactly) and the tracer rows. A one-layer world carries psi (the
cosine bell); an air world carries theta, vapor q and three gases (CO2, N2O, CH4 mole fractions) as one stacked
(5 L, N) tensor, so transport moves all of them in one call with no copies. A tick is one time-split RK3 step of the
layers, one FCT transport step with that step's fluxes, then for a coupled world the air-sea stress (its kinetic
energy back as the air's heat), the land column, and on every `ocean_every`-th tick one ocean step.

Presets: 'w2', 'w5' (Williamson 2 and 5, one layer), 'jet' (one 8 km layer in a balanced 20 cos(lat) m/s jet with
three seeded 100 m bumps, dt 600 s), 'coupled_stack' (three air layers over two ocean layers with a seeded 30% land
mask and random-phase harmonic perturbations, dt 450 s, 15 fast sub-steps, an hourly ocean step), 'held_suarez' (two
air layers under Held & Suarez's forcing), 'aqua_dry' and 'aqua' (three air layers over a slab aquaplanet under the
sun with the radiative column; 'aqua' with vapor, condensation, convection and sea ice; 'aqua5', 'aqua7' the moist
aquaplanet in five and seven layers).

Ticks are captured as HIP graphs per tick kind (air only / with the ocean step) and replayed (torch.cuda.CUDAGraph);
`step(graph=False)` launches every kernel directly. Both give bit-identical results.
"""
import hashlib
import math

import numpy as np
import torch

from . import cases
from . import grid as gridmod
from .constants import CP_DRY, GRAV, KAPPA, PLANET, R_DRY, SOIL_DZ, C_SOIL, K_SOIL
from .gpu import DeviceGrid, csum, dsum
from .layers import DIV_DAMP, Layers
from .ocean import Ocean, visc4
from .transport import BETA, Transport

THIN_WIND = 100.0
GAS_TRACERS = (('co2', 400e-6), ('n2o', 0.32e-6), ('ch4', 1.8e-6))
PRESETS = ('w2', 'w5', 'jet', 'outcrop', 'phillips', 'coupled_stack', 'held_suarez', 'aqua_dry', 'aqua', 'aqua5',
           'aqua7', 'coupled', 'basin')
AQUA = ('aqua_dry', 'aqua', 'aqua5', 'aqua7', 'coupled', 'basin')
# the layered Boussinesq presets' settings (reference: cases.py:208-217): dt, g', del^4's velocity scale, stress
LAYERED = {'outcrop': dict(dt=3600.0, gprime=[0.02], visc4_U=0.5, stress=[0.0, 2.5e-4]),
           'phillips': dict(visc4_U=20.0)}


def _jsonable(x):
    """params as JSON: tuples and arrays become lists, NumPy scalars numbers."""
    if isinstance(x, dict):
        return {str(k): _jsonable(v) for k, v in x.items()}
    if isinstance(x, (list, tuple, np.ndarray)):
        return [_jsonable(v) for v in (x.tolist() if isinstance(x, np.ndarray) else x)]
    if isinstance(x, np.generic):
        return x.item()
    return x


def _mesh_hash(g):
    """A digest of the mesh's numbering and geometry (its cell positions, edge cells and cell areas)."""
    h = hashlib.sha1()
    for a in (g.pos, g.eI, g.eJ, g.areaCell):
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


class World:
    @staticmethod
    def prepare(n=4, preset='w2', params=None):
        """The host-side part of building World(n, preset, params): the planet's mesh, the fast sub-steps' patch
        tables and, for 'coupled_stack', the seeded initial state. NumPy only (nothing touches the GPU), so a serving
        loop can run it on a worker thread; World(n, preset, params, prepared=this) then only uploads and sets up."""
        from .gpu import patch_tables
        P = dict(params or {})
        planet = dict(PLANET, **P.get('planet', {}))
        g = gridmod.for_planet(gridmod.get(n), **planet)
        patch_tables(g)
        out = dict(n=n, preset=preset, params=P, grid=g)
        if preset == 'coupled_stack':
            out['built'] = cases.coupled_stack(g, seed=P.get('seed', 0), land_frac=P.get('land_frac', 0.3))
        return out

    def __init__(self, n=4, preset='w2', params=None, precision=None, state=None, layers=None, prepared=None):
        """A world on level n from case `preset` (params override its settings). layers=K splits a one-layer case
        into K equal layers with g' = 0 (they then move as one); state=(h, b, u) starts a layered case from given
        arrays (reference: world.py:56-58). prepared: World.prepare(n, preset, params)'s result, reused."""
        self.n, self.preset, self.precision = n, preset, precision
        P = self.params = dict(params or {})
        if preset not in PRESETS:
            raise ValueError(f'unknown preset {preset!r}; this build has {PRESETS}')
        self.planet = dict(PLANET, **P.get('planet', {}))
        self.grav = self.planet['grav']
        if prepared is not None and (prepared['n'], prepared['preset'], prepared['params']) != (n, preset, P):
            raise ValueError('prepared for another level, preset or settings')
        self.grid = prepared['grid'] if prepared is not None else gridmod.for_planet(gridmod.get(n), **self.planet)
        self.dg = dg = DeviceGrid(self.grid, precision)
        g = self.grid
        self.air = preset in ('coupled_stack', 'held_suarez') + AQUA
        self.sea = self.physics = None
        self.Ts = self.ice = self.cw = None
        self.rad_every = 0
        if preset in AQUA:
            # three air layers at rest over a slab aquaplanet under the diurnal and seasonal sun, gray radiation (or
            # the k-tables with longwave='ktables'), the bottom layer dragged by the surface; 'aqua' with vapor and
            # sea ice (reference presets 'aqua_dry' and 'aqua', cases.py:225-277): the column step then the exchange
            # of the mass it moves
            from .atmosphere import AirPhysics
            moist = preset != 'aqua_dry'
            coupled = preset in ('coupled', 'basin')
            # with clouds (physics cloud > 0, moist) the cloud water per layer, starting at none and moved as a tracer
            # (reference: world.py:145-148, 294-296)
            cloudy = moist and P.get('physics', {}).get('cloud', 0.0) > 0
            stars = cases.AQUA_N[preset][0] if preset in cases.AQUA_N else [380.0, 305.0]
            if preset in cases.AQUA_N:
                h, b, u, extra = cases.aqua_layers(g, *cases.AQUA_N[preset])
            else:
                h, b, u, extra = cases.aqua_air(g) if moist else cases.aqua_dry(g)
            h = h * (GRAV / self.grav)
            L = self.L = len(stars) + 1
            self.dt = float(P.get('dt', 900.0 * 2.0 ** (5 - n)))
            nu4 = P.get('nu4', visc4(g, P.get('visc4_U', 15.0)))
            floor = P.get('floor', 10.0)
            # (coupled: the air's drag on the surface is the wind stress on the ocean, so its own bottom stress is off)
            self.core = Layers(dg, L, b=b, grav=self.grav, gprime=[0.0] * (L - 1), drag=P.get('drag', 0.0),
                               stress=P.get('stress', [0.0] * L if coupled else
                                            [0.0] * (L - 1) + [cases.SURFACE_DRAG]), nu4=nu4, floor=floor,
                               rotation=P.get('rotation', 1.0), div_damp=P.get('div_damp', DIV_DAMP),
                               air=dict(cp=CP_DRY, kappa=KAPPA, rd=R_DRY), vmix=P.get('vmix', 0.0),
                               thin_wind=P.get('thin_wind', THIN_WIND))
            rows = [extra['theta']] + ([extra['q']] if moist else []) + ([np.zeros_like(extra['q'])] if cloudy else [])
            self.rows = dg.asarray(np.concatenate(rows))
            self.theta = self.rows[0:L]
            if moist:
                self.q = self.rows[L:2 * L]
                self.ice = dg.zeros(g.N)
            if cloudy:
                self.cw = self.rows[2 * L:3 * L]
            self.psi = dg.asarray(np.tile(cases.cosine_bell(g), (L, 1)))      # moved by the exchange only
            if coupled:
                # the ocean under the air (reference presets 'coupled' and 'basin', cases.py:280-287): its mixed layer's
                # temperature is the air's surface temperature; evaporation recalibrated for the ocean's surface
                ocfg = dict(P.get('ocean', {}))
                if preset == 'basin':
                    ocfg.setdefault('walls', (180.0,))
                self.sea = Ocean(dg, ocfg, Tsurf=extra['Ts'], grav=self.grav)
                self.surface_drag = P.get('surface_drag', cases.SURFACE_DRAG)
                ratio = self.sea.cfg['dt'] / self.dt
                if abs(ratio - round(ratio)) > 1e-9 or round(ratio) < 1:
                    raise ValueError(f'the ocean step {self.sea.cfg["dt"]} s must be a whole number of air steps '
                                     f'({self.dt} s)')
                self.ocean_every = int(round(ratio))
                self.Ts = self.sea.T[0]                  # a view: the air's surface physics heats the mixed layer
                self.sea.ice_prev.copy_(self.ice)
            else:
                self.Ts = dg.asarray(extra['Ts'])
            self.tr = Transport(dg, layers=L, beta=P.get('beta', BETA), tracers=len(rows))
            phys = dict(theta_stars=list(stars), floor_=floor, grav=self.grav, moist=1.0 if moist else 0.0)
            if coupled:
                phys.update(k_evap=4.8e-3, firn=2.0)
            if self.grav != GRAV:
                phys.update(mu_ref=1.0e5 / self.grav, mu_bl=1.0e4 / self.grav)
            phys.update(P.get('physics', {}))
            longwave = P.get('longwave', 'gray')
            self.physics = AirPhysics(dg, L, phys, mean_sun=P.get('mean_sun', False), forcing='radiation',
                                      longwave=longwave, lw_tables=P.get('lw_tables'), sw_tables=P.get('sw_tables'))
            if self.physics.lw is not None:
                self.rad_every = max(1, int(round(P.get('rad_dt', 3600.0) / self.dt)))
            self.time = dg.zeros(1)
        elif preset == 'held_suarez':
            # two air layers forced by Held & Suarez's relaxation with the interface crossing (reference preset
            # 'held_suarez', cases.py:680-682): the column step then the exchange of the mass it moves
            from .atmosphere import AirPhysics
            h, b, u, theta = cases.held_suarez(g, seed=P.get('seed', 1))
            h = h * (GRAV / self.grav)
            self.L = 2
            self.dt = float(P.get('dt', 900.0 * 2.0 ** (5 - n)))
            nu4 = P.get('nu4', visc4(g, P.get('visc4_U', 15.0)))
            floor = P.get('floor', 10.0)
            self.core = Layers(dg, 2, b=b, grav=self.grav, gprime=[0.0], drag=P.get('drag', 0.0),
                               stress=P.get('stress', [0.0, cases.SURFACE_DRAG]), nu4=nu4, floor=floor,
                               rotation=P.get('rotation', 1.0), div_damp=P.get('div_damp', DIV_DAMP),
                               air=dict(cp=CP_DRY, kappa=KAPPA, rd=R_DRY), vmix=P.get('vmix', 0.0),
                               thin_wind=P.get('thin_wind', THIN_WIND))
            self.rows = dg.asarray(theta)
            self.theta = self.rows
            self.psi = dg.asarray(np.tile(cases.cosine_bell(g), (2, 1)))      # moved by the exchange only
            self.tr = Transport(dg, layers=2, beta=P.get('beta', BETA), tracers=1)
            phys = dict(theta_star=cases.held_suarez_theta_star(g), floor_=floor, grav=self.grav,
                        **P.get('physics', {}))
            self.physics = AirPhysics(dg, 2, phys, forcing='held_suarez')
            self.time = dg.zeros(1)
        elif self.air:
            built = prepared['built'] if prepared is not None and 'built' in prepared else \
                cases.coupled_stack(g, seed=P.get('seed', 0), land_frac=P.get('land_frac', 0.3))
            h, b, u = built['h'] * (GRAV / self.grav), built['b'], built['u']
            self.L = 3
            self.dt = float(P.get('dt', 900.0 * 2.0 ** (5 - n)))
            nu4 = P.get('nu4', visc4(g, P.get('visc4_U', 15.0)))
            floor = P.get('floor', 10.0)
            self.core = Layers(dg, 3, b=b, grav=self.grav, gprime=[0.0, 0.0], drag=P.get('drag', 0.0),
                               stress=P.get('stress', [0.0, 0.0, 0.0]), nu4=nu4, floor=floor,
                               rotation=P.get('rotation', 1.0), div_damp=P.get('div_damp', DIV_DAMP),
                               air=dict(cp=CP_DRY, kappa=KAPPA, rd=R_DRY), vmix=P.get('vmix', 0.0),
                               thin_wind=P.get('thin_wind', THIN_WIND))
            self.land = built['land'] if P.get('land', True) else None
            gx = [np.full((3, g.N), v) for _, v in GAS_TRACERS]
            self.rows = dg.asarray(np.concatenate([built['theta'], built['q']] + gx))
            self.theta, self.q = self.rows[0:3], self.rows[3:6]
            self.gx = self.rows[6:15].view(3, 3, g.N)
            self.tr = Transport(dg, layers=3, beta=P.get('beta', BETA), tracers=5)
            self.sea = Ocean(dg, dict(gprime=(0.016,), h_layers=(100.0,), **P.get('ocean', {})), Tsurf=built['Ts'],
                             land=self.land, grav=self.grav,
                             soil=dict(dz=SOIL_DZ, c=C_SOIL, k=K_SOIL) if self.land is not None else None)
            self.surface_drag = P.get('surface_drag', cases.SURFACE_DRAG)
            ratio = self.sea.cfg['dt'] / self.dt
            if abs(ratio - round(ratio)) > 1e-9 or round(ratio) < 1:
                raise ValueError(f'the ocean step {self.sea.cfg["dt"]} s must be a whole number of air steps ({self.dt} s)')
            self.ocean_every = int(round(ratio))
        else:
            spec = LAYERED.get(preset, {})
            if preset in ('w2', 'w5'):
                h, b, u = cases.williamson(g, preset)
                h, u = h[None], u[None]
                self.dt = float(P.get('dt', cases.default_dt(n)))
            elif preset == 'outcrop':
                h, b, u = cases.outcrop(g)
                self.dt = float(P.get('dt', spec['dt']))
            elif preset == 'phillips':
                h, b, u, gp3 = cases.phillips(g, 3.0, bump=0.01)
                spec = dict(spec, gprime=[gp3])
                self.dt = float(P.get('dt', 1200.0 * 2.0 ** (5 - n)))
            else:
                h, b, u = cases.jet_bumps(g, seed=P.get('seed', 0))
                self.dt = float(P.get('dt', 600.0))
            if state is not None:
                h, b, u = state[:3]
            h = np.asarray(h, dtype=float).reshape(-1, g.N)
            u = np.asarray(u, dtype=float).reshape(-1, g.E)
            if layers and layers != h.shape[0]:          # (one layer split into equal layers with g' = 0)
                assert h.shape[0] == 1, 'only a one-layer case splits into layers'
                h, u = np.repeat(h / layers, layers, axis=0), np.repeat(u, layers, axis=0)
                P.setdefault('gprime', [0.0] * (layers - 1))
            self.L = h.shape[0]
            U4 = P.get('visc4_U', spec.get('visc4_U'))
            nu4 = P.get('nu4', visc4(g, U4) if U4 is not None else 0.0)
            gp = P.get('gprime', spec.get('gprime', [0.0] * (self.L - 1)))
            self.core = Layers(dg, self.L, b=b, grav=self.grav, gprime=gp, drag=P.get('drag', 0.0),
                               stress=P.get('stress', spec.get('stress', 0.0)), nu4=nu4, floor=P.get('floor', 1.0),
                               rotation=P.get('rotation', 1.0), div_damp=P.get('div_damp', DIV_DAMP))
            self.tracer = P.get('tracer', True)
            self.rows = dg.asarray(np.tile(cases.cosine_bell(g), (self.L, 1)))
            self.psi = self.rows
            self.tr = Transport(dg, layers=self.L, beta=P.get('beta', BETA), tracers=1)
        self.b = self.core.b
        self.h, self.u = dg.asarray(h), dg.asarray(u)
        # the gravity waves' drag on the air (reference: world.py:209-224): the mountains' (the air physics' oro > 0,
        # with the sub-grid orography params['sso'], (4, N) or 'synthetic') and the waves not tied to the ground
        # (gw > 0); both book their heat with the wind stress, so they need the ocean under the air
        self.oro = self.gw = None
        if self.air:
            from .atmosphere import DEFAULTS
            dp = self.physics.p if self.physics is not None else dict(DEFAULTS, grav=self.grav, **P.get('physics', {}))
            if dp['oro'] > 0:
                from .orodrag import OroDrag
                if self.sea is None or P.get('sso') is None:
                    raise ValueError('the mountain drag (oro) needs the coupled ocean and the sub-grid orography (sso)')
                sso = P['sso']
                if isinstance(sso, str):
                    if sso != 'synthetic':
                        raise ValueError("sso: a (4, N) array or 'synthetic'")
                    sso = cases.synthetic_sso(g, seed=P.get('seed', 0), land=getattr(self, 'land', None))
                self.oro = OroDrag(dg, self.L, sso, dp, self.dt)
            if dp['gw'] > 0:
                from .orodrag import WaveDrag
                if self.sea is None:
                    raise ValueError('the wave drag (gw) books its heat with the wind stress: it needs the coupled ocean')
                self.gw = WaveDrag(dg, self.L, dp, self.dt)
                if dp.get('gw_conv', 0.0) > 0:
                    if self.physics is None or getattr(self, 'ice', None) is None:
                        raise ValueError('the waves\' convective source (gw_conv) needs the moist column')
                    self.gw.conv = self.physics.rainp[2]
        # the Moon's and the Sun's tides (reference: world.py:199-208): params['tides'] True / {} for Earth's, or a dict
        # of changes; params 'rough' (N, the floor's roughness h^2) adds the internal tide's drag, 'nbot' its N
        self.tides = None
        tides = P.get('tides')
        if tides is not None and tides is not False:
            if self.sea is None or self.physics is None:
                raise ValueError('tides need the coupled ocean under air')
            from .tides import Tides
            self.tides = Tides(dg, self.sea, self.physics.p, self.planet['radius'], tides, rough=P.get('rough'),
                               nbot=P.get('nbot'))
        self.M = self.h * dg.areaCell[None, :]
        if nu4:
            lam = self.core.lap_max()
            if nu4 * lam * lam * self.dt > 1.0:
                raise ValueError(f'nu4 {nu4:.3g} m^4/s is unstable at dt {self.dt:g} s on n = {n}')
        if self.air:
            self.core.update_gfast(self.h, self.theta)
        self.nfast = int(P.get('nfast', math.ceil(1.2 * self.core.substeps(self.h, self.dt))))
        P['nfast'] = self.nfast
        self.tick = 0
        self._Fm, self._M1 = dg.zeros(self.L, dg.E), dg.zeros(self.L, dg.N)
        self._graphs, self._tick0 = {}, 0
        # the world's own stream (default priority): its ticks queue there, so an ensemble's members overlap
        self._stream = torch.cuda.Stream()
        self._stream.wait_stream(torch.cuda.current_stream())
        # the ocean mixes its mixed layer daily in the reference's coupled worlds (not in coupled_stack, whose ocean
        # steps its layers and tracers only)
        self.sea_mix = self.sea is not None and preset in ('coupled', 'basin')
        # (the mixed layer's heat capacity is derived from its thickness at the last ocean step, and read by the air
        # column and the land column until the next: state)
        sea = () if self.sea is None else (('sea.h', 'sea.u', 'sea.M', 'sea.rows', 'sea.cap') +
                                           (('sea.Tg',) if self.sea.soil is not None else ()) +
                                           (('sea.ice_prev', 'sea.tau_acc', 'sea.q_acc', 'sea.sun_acc', 'sea.fw_acc')
                                            if self.sea_mix else ()))
        self.STATE = ('h', 'u', 'M', 'rows') + sea + (('psi', 'time') if self.physics is not None else ()) + \
            (('Ts',) if self.Ts is not None and not self.sea_mix else ()) + (('ice',) if self.ice is not None else ())

    # ---- the tick ----
    def _kind(self, tick):
        """The tick's kind: bit 0 the ocean's step, bit 2 with the mixed layer's daily mixing, bit 1 the k-tables'
        radiation (every rad_every ticks from the first; reference: world.py:277-284)."""
        k = 0 if self.sea is None or (tick + 1) % self.ocean_every else 1
        if k and self.sea_mix and (tick + 1) % (self.ocean_every * self.sea.mix_every) == 0:
            k |= 4
        if self.rad_every and tick % self.rad_every == 0:
            k |= 2
        return k

    def _tick(self, kind=0):
        """One tick's work on fixed buffers (nothing is allocated), so the sequence can be captured as a graph, in the
        reference's order (world.py:286-358): the layers' step and transport; the wind stress on the ocean with its
        heat; the radiation (the k-tables, saved for the column steps until the next), the column (the radiation or
        Held & Suarez's relaxation, the crossing into wx, the clock), the ground, the ocean's forcing (its sunlight
        below the mixed layer and fresh water), the exchange of the column's mass between the layers; then the ocean's
        step (and its mixing)."""
        self.core.step(self.h, self.u, self.dt, self.nfast, self._Fm, theta=self.theta if self.air else None)
        self.tr.mass_after(self.M, self._Fm, self._M1)
        if self.air or self.tracer:
            self.tr.step_rows(self.M, self._M1, self._Fm, self.rows, order=3)
        self.M.copy_(self._M1)
        sea, moist = self.sea, self.ice is not None
        if sea is not None:
            sea.couple(self.h, self.u, self.surface_drag, self.dt)
            if self.oro is not None:
                self.oro.apply(self.h, self.theta, self.u, self.dt, sea.diss)
            if self.gw is not None:
                self.gw.apply(self.h, self.theta, self.u, self.dt, sea.diss)
            self.core.vmix_heat(self.h, sea.diss, self.theta)
        if self.physics is not None:
            if kind & 2:
                self.physics.radiate(self.h, self.theta, self.Ts, self.time, self.rad_every * self.dt,
                                     q=self.q if moist else None, cw=self.cw)
            self.physics.column(self.h, self.theta, self.Ts, self.u, self.time, self.dt, elev=self.b,
                                q=self.q if moist else None, ice=self.ice if moist else None, cw=self.cw,
                                cap=sea.cap if self.sea_mix else None, qocean=sea.qocean if self.sea_mix else None,
                                qsw=sea.qsw if self.sea_mix else None)
        if sea is not None:
            sea.ground(self.dt)
            if self.sea_mix:
                p = self.physics
                sea.accumulate(self.dt, p.evap if moist else None, p.rain if moist else None,
                               self.ice if moist else None)
        if self.physics is not None:
            self.core.exchange(self.physics.wx, self.h, self.u, self.M, self.psi, theta=self.theta)
        if kind & 1:
            if self.tides is not None:             # (the force at the ocean step's middle)
                self.tides.force(self.time, 0.5 * sea.cfg['dt'])
            sea.step(mix=bool(kind & 4))
            if self.tides is not None:
                self.tides.mix(sea.cfg['dt'])

    def launch_tick(self, graph=True):
        """Queue one tick on the world's stream without waiting for it (step() and Ensemble.step() call this)."""
        kind = self._kind(self.tick)
        gr = self._graphs.get(kind)
        with torch.cuda.stream(self._stream):
            if graph and gr is None and self.tick > self._tick0:
                gr = torch.cuda.CUDAGraph()
                with torch.cuda.graph(gr):
                    self._tick(kind)
                self._graphs[kind] = gr
            if graph and gr is not None:
                gr.replay()
            else:
                self._tick(kind)
        self.tick += 1

    def step(self, k=1, graph=True):
        """k ticks; the first runs directly, later ones replay a captured graph per tick kind."""
        for _ in range(k):
            self.launch_tick(graph)
        self.sync()
        return self

    def sync(self):
        """Wait for the world's queued ticks; later work on the current stream sees their results."""
        self._stream.synchronize()
        torch.cuda.current_stream().wait_stream(self._stream)

    @property
    def ticks_per_hour(self):
        return 3600.0 / self.dt

    # ---- diagnostics ----
    def totals(self):
        """Budgets, each a compensated deterministic sum: per layer volume (or mass) sum h A, M, and each tracer row's
        sum M psi; for a coupled world the ocean's volume, heat content sum M T and salt sum M S per layer."""
        dg, A = self.dg, self.dg.areaCell
        out = {}
        for k in range(self.L):
            out[f'volume{k}'] = csum(dg, self.h[k], A)
            out[f'M{k}'] = csum(dg, self.M[k])
        for y in range(self.rows.shape[0]):
            out[f'tracer{y}'] = csum(dg, self.M[y % self.L], self.rows[y])
        if self.sea is not None:
            s = self.sea
            for k in range(s.L):
                out[f'ocean_volume{k}'] = csum(dg, s.h[k], A)
                out[f'ocean_M{k}'] = csum(dg, s.M[k])
            for y in range(s.rows.shape[0]):
                out[f'ocean_tracer{y}'] = csum(dg, s.M[y % s.L], s.rows[y])
        return out

    def energy(self):
        """Total energy (J): kinetic + potential (+ for air the enthalpy; over a slab its heat c_slab Ts; reference:
        world.py:424-441)."""
        e = self.core.energy(self.h, self.u, self.theta if self.air else None)
        if self.sea_mix:                                 # the ocean's heat in place of the slab's
            e += self.sea.heat()
        elif self.Ts is not None and self.sea is None:
            e += self.physics.p['c_slab'] * csum(self.dg, self.Ts, self.dg.areaCell)
        if self.ice is not None:                         # the vapor's latent energy, the sea ice's -L_f m
            A, p = self.dg.areaCell, self.physics.p
            e += p['Lv'] * sum(csum(self.dg, self.h[k] * self.q[k], A) for k in range(self.L))
            e -= p['Lf'] * csum(self.dg, self.ice, A)
        return e

    def water(self):
        """The air's vapor and cloud water (kg): sum A sum_k mu_k (q_k + cw_k) (reference: world.py:420-422)."""
        if self.ice is None:
            return 0.0
        w = self.q if self.cw is None else self.q + self.cw
        return csum(self.dg, torch.sum(self.h * w, dim=0), self.dg.areaCell)

    def exchange(self, wx):
        """Move thickness wx (L - 1, N; positive up) between neighbouring layers with the tracer and momentum
        (reference: world.py:399-403)."""
        wx = torch.as_tensor(np.asarray(wx, dtype=float), device=self.dg.dev).reshape(self.L - 1, self.dg.N)
        with torch.cuda.stream(self._stream):
            self.core.exchange(wx.contiguous(), self.h, self.u, self.M, self.psi)
        self.sync()

    def fields(self):
        """The cell fields `field` serves for this world (reference: world.py:445-456, for the parts built here)."""
        base = ['h', 'eta'] + ([] if self.air else ['psi'])
        if not self.air:
            return base
        out = base + ['theta', 'ps'] + (['Ts'] if self.Ts is not None else [])
        if self.ice is not None:
            out += ['q', 'ice']
        if self.sea is not None:
            out += ['sst', 'sss']
        return out

    # ---- save / load (the reference's .iw2 files, world.py:581-611; io/save.py) ----
    def save(self, path):
        """Write the world's state, tick and settings to path: the same state gives the same bytes."""
        from . import savefile
        self.sync()
        meta = dict(kind='insilica-world-v2', n=self.n, preset=self.preset, layers=self.L, dt=self.dt, tick=self.tick,
                    params=_jsonable(self.params), precision=self.dg.precision, order=self.grid.order,
                    mesh=_mesh_hash(self.grid))
        arrays = {k: v.detach().cpu().numpy() for k, v in self.state_arrays().items()}
        arrays['b'] = self.b.detach().cpu().numpy()
        savefile.write(path, meta, arrays)

    @classmethod
    def load(cls, path):
        """A world rebuilt from its settings with its saved state copied into the new world's own buffers (so the
        graphs it captures later use them); it continues as the saved world would have."""
        from . import savefile
        meta, arrays = savefile.read(path)
        if meta.get('kind') != 'insilica-world-v2':
            raise ValueError('not a world save file')
        layered = meta['preset'] not in ('coupled_stack', 'held_suarez') + AQUA
        w = cls(meta['n'], meta['preset'], params=meta['params'], precision=meta['precision'],
                state=(arrays['h'], arrays['b'], arrays['u']) if layered else None)
        if meta.get('mesh') != _mesh_hash(w.grid) or meta.get('order') != w.grid.order:
            raise ValueError('the save file was made on another mesh or cell order')
        for k, dst in w.state_arrays().items():
            a = arrays.get(k)
            if a is None or tuple(a.shape) != tuple(dst.shape) or a.dtype != np.float64:
                raise ValueError(f'save file array {k}: expected {tuple(dst.shape)} float64')
            dst.copy_(torch.as_tensor(a, device=w.dg.dev))
        w.tick = w._tick0 = int(meta['tick'])
        torch.cuda.synchronize()
        return w

    def state_arrays(self):
        out = {}
        for k in self.STATE:
            obj = self
            for part in k.split('.'):
                obj = getattr(obj, part)
            out[k] = obj
        return out

    def state_hash(self):
        h = hashlib.sha1()
        for k, v in self.state_arrays().items():
            h.update(v.detach().cpu().numpy().tobytes())
        h.update(str(self.tick).encode())
        return h.hexdigest()

    def field(self, name, layer=0):
        """A cell field (torch float64 on the device)."""
        if name == 'h':
            return self.h[layer]
        if name == 'eta':
            return self.b + torch.sum(self.h, dim=0)
        if name == 'psi' and not self.air:
            return self.psi[layer]
        if name == 'theta' and self.air:
            return self.theta[layer]
        if name == 'q' and self.air:
            return self.q[layer] * 1000.0
        if name == 'Ts' and self.Ts is not None:
            return self.Ts
        if name == 'ice' and self.ice is not None:
            return self.ice
        if name == 'ps' and self.air:
            return self.grav * torch.sum(self.h, dim=0) / 100.0
        if name == 'sst' and self.sea is not None:
            return self.sea.T[0] - 273.15
        if name == 'sss' and self.sea is not None:
            return self.sea.S[0]
        raise KeyError(name)
