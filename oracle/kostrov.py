"""
Kostrov's self-similar circular crack: slip rate in closed form (test
infrastructure: used by the tests only, never by the solver).

Restates cst/kostrov.py:7-48 (Dahlen 1974, eqn 44 for the constant C): a
crack that starts at a point at t = 0 and grows at a constant speed vrup
under a uniform stress drop dtau has, at hypocentral distance r and reduced
time t after the front passes (t = 0 at r / vrup),

    v(r, t) = C dtau / (rho vs) (t + t0) / sqrt(t (t + 2 t0)),   t0 = r / vrup,

    C = b^3 / (I(a^2, b^2) + b^2 / 4 (b + arccos(b) / sqrt(1 - b^2))),
    I(a2, b2) = int_0^inf ((x + b2/2)^2 - x sqrt((x + b2)(x + a2)))
                          / ((x + 1)^2 sqrt(x + b2)) dx,

a = vrup / vp, b = vrup / vs.
"""
import numpy as np


def cee_integral(a2, b2):
    """I(a2, b2) of the module docstring (cst/kostrov.py:7-17)."""
    import scipy.integrate

    def f(x):
        return ((x + 0.5 * b2) ** 2 - x * np.sqrt((x + b2) * (x + a2))) / ((x + 1.0) ** 2 * np.sqrt(x + b2))
    return scipy.integrate.quad(f, 0.0, np.inf)[0]


def cee(a, b):
    """Dahlen's C for a = vrup / vp, b = vrup / vs (cst/kostrov.py:20-29)."""
    b2 = b * b
    return b * b2 / (cee_integral(a * a, b2) + 0.25 * b2 * (b + np.arccos(b) / np.sqrt(1.0 - b2)))


def slip_rate(rho, vp, vs, vrup, dtau, r, t, C=None):
    """Slip rate at hypocentral distance r and reduced time t (cst/kostrov.py:32-48)."""
    t0 = r / vrup
    if C is None:
        C = cee(vrup / vp, vrup / vs)
    return C * dtau / (rho * vs) * (t + t0) / np.sqrt(t * (t + 2.0 * t0))
