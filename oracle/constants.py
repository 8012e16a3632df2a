"""Physical constants of the reference, restated without astropy.

Reference: constants.py:5-12.  astropy defines M_sun = GM_sun / G with the IAU
2015 nominal GM_sun = 1.3271244e20 m^3 s^-2, so G * M_sun is exact and the
kpc^3 / Msun / Gyr^2 value follows from the Julian year and the IAU parsec.
"""
import math

GM_SUN_SI = 1.3271244e20            # m^3 s^-2 (IAU 2015 nominal)
KPC_M = 3.0856775814913673e19       # m (IAU 2012 au, pc = 648000/pi au)
GYR_S = 3.15576e16                  # s (Julian Gyr)

G = GM_SUN_SI * GYR_S ** 2 / KPC_M ** 3          # constants.py:5
KPC_TO_KM = KPC_M / 1e3                          # constants.py:7
GYR_TO_S = GYR_S                                 # constants.py:8
KMS_TO_KPCGYR = GYR_TO_S / KPC_TO_KM             # constants.py:9
KPCGYR_TO_KMS = KPC_TO_KM / GYR_TO_S             # constants.py:10
EPSILON = 1e-12                                  # constants.py:11
TWOPI = 2 * math.pi                              # constants.py:12
