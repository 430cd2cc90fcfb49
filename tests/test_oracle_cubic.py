"""Pins for oracle/cubic.py (and the device Kerr series constants).

solve_1728 (restating CubicEquationSolver.solve, CubicEquationSolver.py:29-90)
is checked against numpy.roots on random cubics from every branch, against
closed forms on the linear/quadratic/triple-root branches, and roots[0] is
checked to be the root the reference's consumer expects (the unique real root
when h > 0, the largest real root when h <= 0).  The physical E-form root is
checked by its residual, its chi3 -> 0 limit and numpy.roots; the
Fuss-Catalan constants the device uses for small z are read from the HIP
source and evaluated here against the closed form.
"""
import math
import os
import re

import numpy as np
import pytest

from oracle.cubic import solve_1728, root0_1728, kerr_root_e

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _match_sets(a, b, tol):
    a = sorted(np.asarray(a, dtype=complex), key=lambda z: (round(z.real, 9), z.imag))
    b = sorted(np.asarray(b, dtype=complex), key=lambda z: (round(z.real, 9), z.imag))
    scale = max(1.0, max(abs(z) for z in b))
    return all(abs(x - y) <= tol * scale for x, y in zip(a, b))


def test_three_real_roots_branch():
    rng = np.random.default_rng(0)
    for _ in range(200):
        r = np.sort(rng.uniform(-5, 5, 3))
        if np.min(np.diff(r)) < 1e-2:
            continue
        a = rng.uniform(0.5, 2)
        b, c, d = -a * r.sum(), a * (r[0] * r[1] + r[0] * r[2] + r[1] * r[2]), -a * r.prod()
        roots = solve_1728(a, b, c, d)
        assert _match_sets(roots, r, 1e-8)
        # roots[0] = 2 j cos(k/3) - b/3a is the largest real root
        assert abs(roots[0] - r.max()) <= 1e-8 * max(1, abs(r.max()))


def test_one_real_root_branch_matches_numpy():
    rng = np.random.default_rng(1)
    for _ in range(200):
        a, b, c, d = rng.uniform(0.5, 2), rng.uniform(-3, 3), rng.uniform(2, 6), rng.uniform(-5, 5)
        if b * b < 3 * a * c:      # monotone cubic -> exactly one real root
            roots = solve_1728(a, b, c, d)
            ref = np.roots([a, b, c, d])
            assert _match_sets(roots, ref, 1e-8)
            real = ref[np.abs(ref.imag) < 1e-9].real
            assert len(real) == 1 and abs(roots[0].real - real[0]) <= 1e-9 * max(1, abs(real[0]))


def test_degenerate_branches_closed_forms():
    # linear: -d/c
    assert solve_1728(0, 0, 4.0, -2.0)[0] == 0.5
    # quadratic: x^2 - 3x + 2 -> (2, 1)
    q = solve_1728(0, 1.0, -3.0, 2.0)
    assert sorted(q.real) == [1.0, 2.0]
    # complex quadratic pair
    q = solve_1728(0, 1.0, 0.0, 1.0)
    assert _match_sets(q, [1j, -1j], 1e-15)
    # triple root (x-2)^3 = x^3 - 6x^2 + 12x - 8: f = g = h = 0 exactly
    assert np.allclose(solve_1728(1.0, -6.0, 12.0, -8.0), 2.0, rtol=1e-15)


def test_root0_cut_and_reference_coefficients():
    # |d| <= 1e-8 is stored as 0 (BaseFDTD11.py:834)
    assert root0_1728(1.0, 1.0, 1.0, 1e-9) == 0.0
    # reference-shaped cubic: (alpha3 chi3)^2 x^3 + 2 Re(alpha3 eps chi3) x^2 + |eps|^2 x - |D/eps0|^2
    a3 = float(np.float32(0.7))
    w0 = 2 * math.pi * 20e9
    wp2 = 1.5 * w0 ** 2
    w = 2 * math.pi * 9e9
    eps = 1 + wp2 / (w0 ** 2 - w ** 2 + 1j * 0.1 * w0 * w)
    a, b, c = (a3 * 1e-3) ** 2, 2 * np.real(a3 * eps * 1e-3), abs(eps) ** 2
    # Cardano's S+U cancels when a x^3 << c x: the absolute error of x is ~eps*|S|
    # (|S| ~ sqrt(c/3a) ~ 2e3 here), a floor the reference's algorithm carries.
    floor = 1e-10
    for dd in (1e-6, 1e-2, 1.0, 1e3, 1e6):
        x = root0_1728(a, b, c, -dd)
        res = a * x ** 3 + b * x ** 2 + c * x - dd
        assert abs(res) <= 1e-9 * dd + floor
        assert x > 0
        # the physical meaning: x = |E|^2 with |eps + alpha3 chi3 x|^2 x = |D/eps0|^2
        assert abs(abs(eps + a3 * 1e-3 * x) ** 2 * x - dd) <= 1e-9 * dd + floor


@pytest.mark.parametrize("chi3", [1e-19, 1e-6, 1e-2, 1.0])
def test_kerr_root_e_residual_and_numpy(chi3):
    rng = np.random.default_rng(2)
    for _ in range(100):
        eps = rng.uniform(1, 3)
        u = rng.normal() * (1.8e8 if chi3 < 1e-10 else 3.0)
        E = kerr_root_e(u, eps, chi3)
        assert abs(chi3 * E ** 3 + eps * E - u) <= 1e-13 * (abs(u) + abs(eps * E))
        r = np.roots([chi3, 0.0, eps, -u])
        real = r[np.abs(r.imag) <= 1e-9 * np.abs(r)].real
        assert len(real) == 1 and abs(real[0] - E) <= 1e-10 * abs(E)


def test_kerr_root_e_linear_limit():
    for u in (-3.0, 1e-12, 2.5e8):
        assert kerr_root_e(u, 2.0, 0.0) == u / 2.0
        assert abs(kerr_root_e(u, 2.0, 1e-30) - u / 2.0) <= 1e-14 * abs(u)


def _device_series_coefficients():
    src = open(os.path.join(ROOT, "py-fdtd_pic_amd", "csrc", "device_common.hpp")).read()
    body = src[src.index("kerr_root_e"):]
    body = body[:body.index("return x * y;")]
    first = float(re.search(r"double y = ([0-9.]+);", body).group(1))
    rest = [float(v) for v in re.findall(r"fma\(y, -z, ([0-9.]+)\)", body)]
    return [first] + rest      # highest order first


def test_device_series_constants_are_fuss_catalan():
    coeffs = _device_series_coefficients()
    k = len(coeffs) - 1
    expect = [math.comb(3 * j, j) / (2 * j + 1) for j in range(k, -1, -1)]
    assert coeffs == expect


def test_device_series_matches_closed_form_below_cut():
    coeffs = _device_series_coefficients()
    for z in np.linspace(0.0, 0.0099, 40):
        y = 0.0
        for c in coeffs:
            y = y * (-z) + c
        # closed form with x = 1, eps = 1, chi3 = z: E = y
        ref = kerr_root_e(1.0, 1.0, z) if z > 0 else 1.0
        assert abs(y - ref) <= 2e-15 * abs(ref)   # a few ulp: sinh/asinh rounding
