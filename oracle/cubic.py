"""Oracle: cubic root solvers used on the Kerr path.  TEST INFRASTRUCTURE ONLY.

`solve_1728` restates the published 1728.org closed form as the reference
vendors it (CubicEquationSolver.py:29-105): linear, quadratic, triple-root,
three-real (trigonometric) and one-real (Cardano) branches, in that order,
with the same branch tests (`a == 0 and b == 0`, `a == 0`,
`f == 0 and g == 0 and h == 0`, `h <= 0`, `h > 0`).  The reference keeps
roots as a numpy array; only roots[0] is ever consumed
(BaseFDTD11.py:837-838), so `root0_1728` returns that element.

`kerr_root_e` is the E-form solve used by the physical model: the real root
of chi3*E^3 + eps*E - u = 0 that continues the linear solution u/eps.  It is
the build's own formulation (the reference's x=|E|^2 form cancels
catastrophically at chi3*E^2 ~ 1e-3), pinned in the tests by the residual,
by the chi3 -> 0 limit and by numpy.roots on small cases.
"""
import math

import numpy as np


def _find_f(a, b, c):
    # CubicEquationSolver.py:94-95
    return ((3.0 * c / a) - ((b ** 2.0) / (a ** 2.0))) / 3.0


def _find_g(a, b, c, d):
    # CubicEquationSolver.py:99-100
    return (((2.0 * (b ** 3.0)) / (a ** 3.0)) - ((9.0 * b * c) / (a ** 2.0)) + (27.0 * d / a)) / 27.0


def _find_h(g, f):
    # CubicEquationSolver.py:104-105
    return (g ** 2.0) / 4.0 + (f ** 3.0) / 27.0


def solve_1728(a, b, c, d):
    """All roots of a x^3 + b x^2 + c x + d (CubicEquationSolver.py:29-90)."""
    if a == 0 and b == 0:                       # :30-31 linear
        return np.array([(-d * 1.0) / c])
    if a == 0:                                  # :33-45 quadratic
        disc = c * c - 4.0 * b * d
        if disc >= 0:
            r = math.sqrt(disc)
            return np.array([(-c + r) / (2.0 * b), (-c - r) / (2.0 * b)])
        r = math.sqrt(-disc)
        return np.array([(-c + r * 1j) / (2.0 * b), (-c - r * 1j) / (2.0 * b)])
    f = _find_f(a, b, c)
    g = _find_g(a, b, c, d)
    h = _find_h(g, f)
    if f == 0 and g == 0 and h == 0:            # :51-56 triple root
        if (d / a) >= 0:
            x = (d / (1.0 * a)) ** (1 / 3.0) * -1
        else:
            x = (-d / (1.0 * a)) ** (1 / 3.0)
        return np.array([x, x, x])
    if h <= 0:                                  # :58-72 three real roots
        i = math.sqrt(((g ** 2.0) / 4.0) - h)
        j = i ** (1 / 3.0)
        k = math.acos(-(g / (2 * i)))
        L = j * -1
        M = math.cos(k / 3.0)
        N = math.sqrt(3) * math.sin(k / 3.0)
        P = (b / (3.0 * a)) * -1
        x1 = 2 * j * math.cos(k / 3.0) - (b / (3.0 * a))
        return np.array([x1, L * (M + N) + P, L * (M - N) + P])
    # h > 0: one real root, Cardano (:74-90)
    R = -(g / 2.0) + math.sqrt(h)
    S = R ** (1 / 3.0) if R >= 0 else (-R) ** (1 / 3.0) * -1
    T = -(g / 2.0) - math.sqrt(h)
    U = (T ** (1 / 3.0)) if T >= 0 else ((-T) ** (1 / 3.0)) * -1
    x1 = (S + U) - (b / (3.0 * a))
    x2 = -(S + U) / 2 - (b / (3.0 * a)) + (S - U) * math.sqrt(3) * 0.5j
    x3 = -(S + U) / 2 - (b / (3.0 * a)) - (S - U) * math.sqrt(3) * 0.5j
    return np.array([x1, x2, x3])


def root0_1728(a, b, c, d):
    """Re(roots[0]) as consumed by Nonlin_Cubic_Solver (BaseFDTD11.py:834-838).

    The reference stores |d| <= 1e-8 cells as 0 (`out = np.zeros`, :825,:834).
    """
    if abs(d) > 1e-8:
        return float(np.real(solve_1728(a, b, c, d)[0]))
    return 0.0


# ---------------------------------------------------------------------------
# Physical model: E-form Kerr root
# ---------------------------------------------------------------------------
def kerr_root_e(u, eps, chi3):
    """Real root E of chi3*E^3 + eps*E = u that continues E = u/eps (FP64).

    Reference-free definition (the reference has no E-form solver); checked
    against numpy.roots and the residual in tests/test_oracle_cubic.py.
    """
    u = float(u)
    if chi3 == 0.0 or u == 0.0:
        return u / eps
    # depressed monic form E^3 + p E + q = 0
    p = eps / chi3
    q = -u / chi3
    if p > 0:
        # one real root: E = -2 sqrt(p/3) sinh(asinh(3q/(2p) sqrt(3/p)) / 3)
        s = math.sqrt(p / 3.0)
        arg = (1.5 * q / p) / s
        return -2.0 * s * math.sinh(math.asinh(arg) / 3.0)
    # p < 0 (defocusing): pick the root nearest the linear solution
    roots = np.roots([1.0, 0.0, p, q])
    real = roots[np.abs(roots.imag) <= 1e-12 * np.maximum(1.0, np.abs(roots.real))].real
    lin = u / eps
    return float(real[np.argmin(np.abs(real - lin))])
