"""Water moving over and through the land, and the ice sheets' flow (reference: src/insilica/land.py:1-60,
kernels/land.cu): the reference's `LandWater` on the HIP library (isw_land_route, isw_ice_flow).

Surface water runs toward lower water surfaces at a set speed whatever the grid, collecting in lakes; soil water seeps
between water tables; what reaches the sea waits in `mouth` (kg/m^2 of the sea cell) for the ocean's fresh water. The
ice sheets flow by Glen's shallow-ice law; ice reaching the sea becomes that cell's sea ice (`calved`, and the ocean's
`ice_prev` when set). The columns' side of the land's water (rain on the ponds, soaking in, springs, evaporation,
freezing) is the air column's.
"""
import torch

RIVER_SPEED = 1.0                       # m/s, the surface flow's speed
K_SEEP = 2.0e4 / 3600.0                 # m^2/s, the soil's seepage coefficient
RHO_W = 1000.0                          # kg/m^3
GLEN_A = 1.0e-16 / (365.0 * 86400.0)    # Pa^-3 s^-1, Glen's flow law at about -10 C
GRAV = 9.81                             # m/s^2, the ice flow's own gravity on Earth


class LandWater:
    def __init__(self, dg, b, land, soilcap, pond, soil, ice, rho_ice=917.0, uflow=RIVER_SPEED, kseep=K_SEEP,
                 grav=GRAV):
        """b, land, soilcap: terrain (m), mask (> 0.5 land) and the bucket's capacity; pond, soil, ice: the world's
        buffers (kg/m^2, device tensors), changed in place (reference: land.py:17-31)."""
        from . import _hip
        _hip.lib()
        self.dg, self.b, self.land, self.soilcap = dg, b, land, soilcap
        self.pond, self.soil, self.ice = pond, soil, ice
        self.rho_ice, self.uflow, self.kseep = rho_ice, uflow, kseep
        N = dg.N
        z = lambda: torch.zeros(N, dtype=torch.float64, device=dg.dev)
        self.mouth, self.calved = z(), z()               # kg/m^2 of each sea cell: water waiting for the ocean; ice given
        self._work = torch.zeros(4 * N, dtype=torch.float64, device=dg.dev)
        self.ice_prev = None                             # the ocean's sea-ice record, when there is one
        self.geo = None                                  # (the geology's accumulators are not in this build)
        self.gam = 2.0 * GLEN_A * (rho_ice * grav) ** 3 / 5.0

    def step(self, dt):
        """One routing step of dt seconds, then the ice's (reference: land.py:33-46)."""
        from . import _hip
        if self.geo is not None:
            raise NotImplementedError('the geology\'s accumulators are not in this build')
        P = _hip.ptr
        _hip.call('isw_land_route', self.dg.ref, P(self.b), P(self.land), P(self.soilcap), P(self.pond), P(self.soil),
                  P(self.ice), RHO_W, float(self.rho_ice), float(self.uflow * dt), float(self.kseep * dt),
                  P(self._work), P(self.mouth), None, _hip.stream())
        self.ice_step(dt)

    def ice_step(self, dt):
        """The ice sheets' flow over dt seconds (reference: land.py:48-60)."""
        from . import _hip
        P = _hip.ptr
        _hip.call('isw_ice_flow', self.dg.ref, P(self.b), P(self.land), P(self.ice), float(self.rho_ice),
                  float(self.gam * dt), P(self._work), P(self.ice_prev), P(self.calved), None, _hip.stream())
