"""The oracle's world tick -- TEST INFRASTRUCTURE ONLY.

Restates the tick order of the reference's World._tick (src/insilica/world.py:286-358) for the parts this build
covers (the layered core, FCT transport, the air-sea stress with its heat, the land column, the hourly ocean step with
its convective adjustment and, in the coupled worlds, the ocean's forcing from the air column and its daily mixing,
the Held-Suarez column and the gray radiative air column, dry or moist, with or without clouds) and the Ocean.step
order (src/insilica/ocean.py:226-295), on the CPU oracle. Built from a
product World's settings and starting state, so that the two can be stepped side by side.
"""
import numpy as np

from . import Grid, Layers, couple_stress, land_column, mass_update, ocn_cap, transport, vmix_heat


def _np(t):
    return np.ascontiguousarray(t.detach().cpu().numpy()).copy()


class OracleWorld:
    def __init__(self, w):
        """Mirror of the product World w (its grid, settings and current state)."""
        self.og = og = Grid(w.grid)
        self.L, self.dt, self.nfast, self.air, self.ocean_every = w.L, w.dt, w.nfast, w.air, getattr(w, 'ocean_every', 0)
        c = w.core
        self.core = Layers(og, c.L, b=_np(c.b), grav=c.grav, gprime=c.gprime or (), drag=c.drag, stress=c.stress,
                           nu4=c.nu4, floor=c.floor, rotation=1.0, div_damp=c.div_damp,
                           air=dict(cp=c.cpa, kappa=c.kap, rd=c.rda) if c.air else None, vmix=c.vmix, thin=c.thin,
                           wall=_np(c.wall) if c.wall is not None else None, thin_wind=c.thin_wind, shelf=c.shelf,
                           vmix_cd=c.vmix_cd)
        self.core.fV[:] = _np(c.fV)
        self.core.Gc[:] = _np(c.Gc)
        self.h, self.u, self.M, self.rows = _np(w.h), _np(w.u), _np(w.M), _np(w.rows)
        self.moved = w.air or getattr(w, 'tracer', True)
        self.beta = w.tr.beta
        self.tick = w.tick
        self.Fm, self.M1 = np.zeros((self.L, og.E)), np.zeros((self.L, og.N))
        self.sea = None
        self.physics = None
        self.Ts = self.ice = None
        if getattr(w, 'physics', None) is not None:                 # the Held-Suarez or the dry radiative column
            p = w.physics
            if p.forcing == 'held_suarez':
                self.physics = dict(kind='hs', par=np.array(p.par), hs=np.array(p.hs), lat=_np(p.lat))
            else:
                if p.lw is not None or p.perot is not None or p.sfc:
                    raise NotImplementedError('the oracle world steps the gray dry column with the edge-wind speed')
                self.physics = dict(kind='air', par=np.array(p.par), mode=p.mode, pos=_np(p.pos), grid=w.grid,
                                    insol=_np(p.insol) if p.insol is not None else None, b=_np(w.b))
                self.Ts = _np(w.Ts)
                self.ice = _np(w.ice) if w.ice is not None else None
                self.cloudy = getattr(w, 'cw', None) is not None      # (the cloud water: rows 2L..3L)
            self.psi, self.time = _np(w.psi), float(w.time[0])
        # the drags (the product World's settings and tables: Perot's coefficients, the edges' normals, the sub-grid
        # orography, the settings in the kernels' order)
        self.oro = self.gw = None
        if getattr(w, 'oro', None) is not None:
            o = w.oro
            self.oro = dict(op=_np(o.op), nrm=_np(o.nrm), sso=_np(o.sso),
                            par=np.array([getattr(o.par, k) for k, _ in o.par._fields_]))
        if getattr(w, 'gw', None) is not None:
            o = w.gw
            if o.cmean is not None:
                raise NotImplementedError('the oracle world steps the waves without the convective source')
            self.gw = dict(op=_np(o.op), nrm=_np(o.nrm), par=np.array([getattr(o.par, k) for k, _ in o.par._fields_]))
        if w.sea is not None:
            s = w.sea
            self.surface_drag = w.surface_drag
            self.rho, self.cw, self.c_land, self.land_drag = s.cfg['rho'], s.cfg['cw'], s.c_land, s.cfg['land_drag']
            sc = s.core
            self.score = Layers(og, sc.L, b=_np(sc.b), grav=sc.grav, gprime=sc.gprime, drag=sc.drag, stress=sc.stress,
                                nu4=sc.nu4, floor=sc.floor, div_damp=sc.div_damp, vmix=sc.vmix, thin=sc.thin,
                                wall=_np(sc.wall) if sc.wall is not None else None, shelf=sc.shelf, vmix_cd=sc.vmix_cd)
            self.score.fV[:] = _np(sc.fV)
            self.score.Gc[:] = _np(sc.Gc)
            self.sea = dict(h=_np(s.h), u=_np(s.u), M=_np(s.M), rows=_np(s.rows), cap=_np(s.cap),
                            tau_acc=_np(s.tau_acc), diss=_np(s.diss), dt=s.cfg['dt'], nfast=s.nfast, L=s.L,
                            land=_np(s.land) if s.land is not None else None,
                            Tg=_np(s.Tg) if s.soil is not None else None,
                            dz=np.asarray(s.soil['dz'], dtype=float) if s.soil is not None else None,
                            c_soil=s.soil['c'] if s.soil is not None else 0.0,
                            k_soil=s.soil['k'] if s.soil is not None else 0.0,
                            Fm=np.zeros((s.L, og.E)), M1=np.zeros((s.L, og.N)),
                            eos=(s.cfg['alpha'], s.cfg['alpha2'], s.cfg['beta']), c_bot=s.cfg['bottom_stress'])
            self.sea_mix = bool(getattr(w, 'sea_mix', False))
            # the tides (the product World's settings): the sky, the cells' positions, the potential the ocean's
            # layers read, the internal tide's drag per edge
            self.tides = None
            if getattr(w, 'tides', None) is not None:
                from insilica.tides import sky_array
                t = w.tides
                self.tides = dict(tp=sky_array(t.sky), pos=np.ascontiguousarray(w.grid.pos, dtype=float),
                                  V=np.zeros(og.N), dsum=_np(t.dsum), cfg=dict(t.cfg),
                                  h2=_np(t.h2) if t.h2 is not None else None,
                                  nbot=_np(t.nbot) if t.nbot is not None else None,
                                  cbot=np.zeros(og.E) if t.h2 is not None else None)
                self.score.set_tides(self.tides['V'], self.tides['cbot'], t.cfg['cd'])
            if self.sea_mix:                       # the coupled worlds' mixed layer: its forcing sums, its settings
                from insilica.ocean import ALPHA_REF
                c = s.cfg
                par = {k: c[k] for k in ('S_ref', 'hmin', 'hmax', 'dTmin', 'm_kt', 'n_kt', 'sw_R', 'sw_z1', 'sw_z2',
                                         'kappa_d', 'dwf', 'ri_b', 'reach', 'h_ent', 'det_class', 'fw_local', 'alpha',
                                         'alpha2', 'beta', 'kappa_h', 'mix_dt')}
                par.update(grav=sc.grav, alpha_ref=ALPHA_REF, rho_w=c['rho'], cw=c['cw'])
                for k in ('ice_prev', 'q_acc', 'sun_acc', 'fw_acc'):
                    self.sea[k] = _np(getattr(s, k))
                self.sea.update(par=par, mix_every=s.mix_every, taum=np.zeros(og.N), wx=np.zeros((s.L - 1, og.N)),
                                qocean=np.zeros(og.N), qsw=np.zeros(og.N))
                if par['ri_b'] > 0 or par['kappa_h'] > 0:
                    raise NotImplementedError('the oracle world mixes without ri_b and kappa_h')
                self.Ts = self.sea['rows'][0]      # a view: the mixed layer is the air's surface

    def _kind(self, tick):
        """0 air only; 1 with the ocean's step; 5 with its step and the mixed layer's mixing."""
        if self.sea is None or (tick + 1) % self.ocean_every:
            return 0
        if getattr(self, 'sea_mix', False) and (tick + 1) % (self.ocean_every * self.sea['mix_every']) == 0:
            return 5
        return 1

    def step(self, k=1):
        for _ in range(k):
            self._tick(self._kind(self.tick))
            self.tick += 1
        return self

    def _tick(self, kind):
        """The reference's tick order (world.py:286-358): the layers and transport; the wind stress and its heat; the
        column; the ground; the ocean's forcing; the exchange of the column's mass; the ocean's step and mixing."""
        og, L = self.og, self.L
        theta = self.rows[0:L] if self.air else None   # a view: the air's theta is the first L rows
        self.core.step(self.h, self.u, self.dt, self.nfast, self.Fm, theta=theta)
        mass_update(og, L, self.M, self.Fm, self.M1)
        if self.moved:
            transport(og, L, self.M, self.M1, self.Fm, self.rows, beta=self.beta, order=3)
        self.M[...] = self.M1
        s = self.sea
        mix = s is not None and getattr(self, 'sea_mix', False)
        if s is not None:
            couple_stress(og, L, self.h, self.u, s['h'], s['u'], self.surface_drag, self.rho, self.dt, s['diss'],
                          s['tau_acc'], wall=_wall(self.score), land_drag=self.land_drag)
            if self.oro is not None or self.gw is not None:
                from .drag import oro_drag, wave_drag
                if self.oro is not None:
                    d = self.oro
                    self.u, s['diss'], _, _ = oro_drag(og, self.h, theta, self.u, d['op'], d['nrm'], d['sso'],
                                                       d['par'], self.dt, s['diss'])
                if self.gw is not None:
                    d = self.gw
                    self.u, s['diss'], _ = wave_drag(og, self.h, theta, self.u, d['op'], d['nrm'], d['par'], self.dt,
                                                     s['diss'])
            vmix_heat(og, L, self.h, s['diss'], theta, grav=self.core.grav, cp=self.core.cpa, kappa=self.core.kap)
        wx = o = None
        if self.physics is not None:
            from .column import air_column, cell_speed, hs_column
            ph = self.physics
            if ph['kind'] == 'hs':
                th, wx, _, _ = hs_column(ph['par'], ph['hs'], ph['lat'], self.h, self.rows[0:L], self.dt)
            else:
                uabs = cell_speed(ph['grid'], self.u[L - 1])
                moist = self.ice is not None
                cloudy = getattr(self, 'cloudy', False)
                o = air_column(ph['par'], ph['mode'], np.array([self.time]), ph['pos'], uabs, self.h, self.rows[0:L],
                               self.Ts, self.dt, insol=ph['insol'], elev=ph['b'],
                               q=self.rows[L:2 * L] if moist else None, ice=self.ice,
                               cw=self.rows[2 * L:3 * L] if cloudy else None, cap=s['cap'] if mix else None,
                               ocean=mix)
                th, wx = o['theta'], o['wx']
                self.Ts[...] = o['Ts']                 # (in place: with an ocean, the mixed layer's row)
                if moist:
                    self.rows[L:2 * L], self.ice = o['q'], o['ice']
                if cloudy:
                    self.rows[2 * L:3 * L] = o['cw']
                if mix:
                    s['qocean'][...], s['qsw'][...] = o['qocean'], o['qsw']
            self.rows[0:L] = th
            self.time += self.dt
        if s is not None:
            if s['Tg'] is not None and s['land'] is not None:
                land_column(og, len(s['dz']), s['land'], s['rows'], s['cap'], s['Tg'], s['dz'], s['c_soil'],
                            s['k_soil'], self.dt)
            if mix:
                from . import ocean as oo
                Lo, N, p = s['L'], og.N, s['par']
                if self.ice is not None:
                    oo.fresh(N, p, s['h'][0], s['rows'][Lo], o['evap'], o['rain'], self.ice, s['ice_prev'], self.dt,
                             s['fw_acc'], land=s['land'], hw=s['h'][0], M=s['M'][0], area=og.arr['areaCell'])
                oo.pen(N, Lo, p, s['h'], s['rows'][:Lo], s['qocean'], s['qsw'], self.dt, s['q_acc'], s['sun_acc'],
                       land=s['land'])
        if self.physics is not None:
            self.core.exchange(wx[:L - 1], self.h, self.u, self.M, self.psi, theta=self.rows[0:L])
        if kind & 1:
            from .ocean import overturn
            Lo, N = s['L'], og.N
            td = getattr(self, 'tides', None)
            if td is not None:
                from .tides import tide_itd, tide_pot
                tide_pot(og, td['tp'], self.time, 0.5 * s['dt'], td['pos'], td['V'])
                if td['h2'] is not None:
                    tide_itd(og, Lo, s['h'], s['rows'][:Lo], s['rows'][Lo:], td['h2'], td['cfg']['kappa_it'],
                             (self.score.grav,) + tuple(s['eos']), td['cbot'], td['nbot'])
            overturn(N, Lo, s['h'], s['rows'][:Lo], s['rows'][Lo:], self.score.grav, s['eos'][0], s['eos'][1],
                     s['eos'][2], land=s['land'])
            self.score.step(s['h'], s['u'], s['dt'], s['nfast'], s['Fm'])
            mass_update(og, s['L'], s['M'], s['Fm'], s['M1'])
            transport(og, s['L'], s['M'], s['M1'], s['Fm'], s['rows'], order=3)
            s['M'][...] = s['M1']
            ocn_cap(og, s['h'], self.rho * self.cw, s['cap'], s['land'], self.c_land)
            if kind & 4:
                self._mix()
            if td is not None:
                from .tides import tide_mix
                tide_mix(og, Lo, s['h'], s['u'], s['rows'][:Lo], s['rows'][Lo:], s['land'], s['c_bot'], self.rho,
                         td['cfg']['gamma'], s['dt'], (self.score.grav,) + tuple(s['eos']), td['dsum'],
                         td['cfg']['cd'], td['cbot'], td['cfg']['zeta'])

    def _mix(self):
        """Ocean.step's mixing (ocean.py:257-295): the interval's forcing, the mixed layer's column step, salt with the
        water, the exchange between the ocean's layers, the sums cleared."""
        from . import ocean as oo
        og, s = self.og, self.sea
        Lo, N, p = s['L'], og.N, s['par']
        dt = p['mix_dt']
        g = og.host
        tau = np.sqrt(s['tau_acc'] * (1.0 / dt))
        s['tau_acc'][...] = tau
        ce = g.cellEdge
        ke = np.zeros(N)
        for m in range(6):                            # (the cell speed of the rms stress, in table order)
            e = ce[:, m]
            ok = e >= 0
            ee = np.where(ok, e, 0)
            ke = ke + np.where(ok, g.le[ee] * g.de[ee] * 0.25 * tau[ee] * tau[ee], 0.0)
        s['taum'][...] = np.sqrt(2 * ke / g.areaCell)
        for k in ('q_acc', 'sun_acc', 'fw_acc'):
            s[k] /= dt
        T, S = s['rows'][:Lo], s['rows'][Lo:]
        oo.column(N, Lo, p, s['h'], T, S, s['taum'], s['q_acc'], s['fw_acc'], s['sun_acc'], dt, s['wx'],
                  land=s['land'])
        seq = p['reach'] > 0
        oo.xch_salt(N, Lo, s['wx'], s['h'], S, seq=seq)
        self.score.exchange(s['wx'], s['h'], s['u'], s['M'], T, seq=seq)
        for k in ('tau_acc', 'q_acc', 'sun_acc', 'fw_acc'):
            s[k][...] = 0.0
        ocn_cap(og, s['h'], self.rho * self.cw, s['cap'], s['land'], self.c_land)

    def state(self):
        out = dict(h=self.h, u=self.u, M=self.M, rows=self.rows)
        if self.physics is not None:
            out.update(psi=self.psi, time=np.array([self.time]))
        if self.Ts is not None:
            out['Ts'] = self.Ts
        if getattr(self, 'ice', None) is not None:
            out['ice'] = self.ice
        if self.sea is not None:
            # (the forcing sums are state only where the mixed layer mixes, as in the product's STATE)
            mixed = ('ice_prev', 'tau_acc', 'q_acc', 'sun_acc', 'fw_acc') if getattr(self, 'sea_mix', False) else ()
            for k in ('h', 'u', 'M', 'rows', 'Tg', 'cap') + mixed:
                if self.sea.get(k) is not None:
                    out['sea.' + k] = self.sea[k]
            if getattr(self, 'sea_mix', False):
                out.pop('Ts', None)                    # (the mixed layer's row)
        return out


def _wall(layers):
    return layers.wall
