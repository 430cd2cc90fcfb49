"""The reference's only published result, measured reflection against
Fresnel (MasterController.py:533-569, BaseFDTD11.py:882-923, "Reflection vs
frequency.png"), reproduced with the oracle: the two-pass reflection of a
Lorentz half-space, measured with the compiled FP64 oracle
(oracle/fdtd_cpu.c, bitwise oracle/physical.py), equals the closed-form
reflection of the same interface on the Yee grid (refcases.discrete_fresnel)
to 2e-6 wherever the pulse carries power (246 of 256 frequencies, 105-300
THz), and the continuum Fresnel of the analytic Lorentz eps to 2e-2 (the
grid's dispersion at 40 cells per wavelength).

Catches: a wrong ADE coefficient (A, B or C: eps_num moves; the planted
test shifts C by 1 % and misses by > 1e-3), E formed without eps_inf, or a
D update with the wrong Courant factor (the slab's impedance moves).
"""
import numpy as np

from refcases import discrete_fresnel, fresnel_spec

FREQS = np.linspace(100e12, 300e12, 256)


def _measured(p, mutate=None):
    from fdtd_pic_amd import sweep
    from oracle import cpu_ref, physical as ph
    X = []
    for with_slab in (False, True):
        m = ph.build(sweep.pass_params(p, with_slab))
        if mutate and with_slab:
            mutate(m)
        sim, _ = cpu_ref.run_model(m, threads=8)
        X.append(ph.dft(sim.probe_out, m["dt"], FREQS)[0])
    band = np.abs(X[0]) > 0.1 * np.abs(X[0]).max()
    return np.abs(X[1] - X[0]) / np.abs(X[0]), band


def test_half_space_reflection_is_discrete_fresnel():
    p = fresnel_spec()
    R, band = _measured(p)
    r_num, r_cont = discrete_fresnel(p, FREQS)
    assert band.sum() >= 240
    assert np.abs(R - r_num)[band].max() < 2e-6
    assert np.abs(R - r_cont)[band].max() < 2e-2


def test_planted_ade_coefficient_fails_fresnel_pin():
    p = fresnel_spec()

    def wrong_C(m):                  # C of the ADE off by 1 %
        m["C"] = m["C"] * 1.01
    R, band = _measured(p, wrong_C)
    r_num, _ = discrete_fresnel(p, FREQS)
    assert np.abs(R - r_num)[band].max() > 1e-3
