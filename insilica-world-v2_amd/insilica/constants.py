"""Physical constants (SI). The values are the reference's (src/insilica/constants.py:4-14), so that a world here
and there start from the same numbers."""
import math

R_PLANET = 6.371e6            # m
OMEGA = 2 * math.pi / 86400   # rad/s: a 24-hour day
GRAV = 9.80616                # m/s^2 (Williamson et al. 1992)
T_REF = 273.15                # K
R_DRY = 287.04                # J/kg/K
CP_DRY = 1004.6               # J/kg/K
KAPPA = R_DRY / CP_DRY
P0 = 1.0e5                    # Pa, potential temperature's reference pressure
PLANET = dict(radius=R_PLANET, omega=OMEGA, grav=GRAV)
C_SOIL, K_SOIL = 2.0e6, 1.0   # soil volumetric heat capacity (J/m^3/K) and conductivity (W/m/K)
SOIL_DZ = (0.1, 0.2, 0.4, 0.8, 1.6, 3.2)
