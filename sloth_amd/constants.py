"""Physical constants used on the traced path.

Values equal scipy.constants (CODATA 2022), which the reference imports
(src/core/propagator.py:14-15, :145); tests check the equality.
"""
C_LIGHT = 299792458.0          # speed of light [m/s]
E_CHARGE = 1.602176634e-19     # elementary charge [C]
EPSILON_0 = 8.8541878188e-12   # vacuum permittivity [F/m]
# m_e eps0 / e^2 as hard-coded by the reference (src/core/propagator.py:48-50, :73, :88):
# critical density n_cr = ALPHA_CRIT * omega^2.
ALPHA_CRIT = 3.14207787e-4
PI = 3.141592653589793
