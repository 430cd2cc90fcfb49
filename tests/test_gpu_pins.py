"""Closed-form pins of the device path, independent of the oracle.

* The deposit of a cold beam drifting transversely at v_x sums to
  J' = Z0 n q v_x L (its surface current times Z0): the device's current
  units follow from density, extent and drift, not from a restated formula.
* Heavy particles (they do not move) with that drift form a current sheet
  that radiates E = -J'/2 on both sides and H' = -/+J'/2 left/right (the
  jump condition [H'] = J'); the sheet switches on at step 0, so the field
  at +-400 cells rings ~0.3 % around that value.
* BaseFDTD11.CPML_Psi_e_Update / CPML_Psi_m_Update run on the device
  (fdtd_step_op) with the difference frozen: psi_n = c d (1 - b^n)/(1 - b)
  and the field moves by -/+ Cb c d sum_k (1 - b^k)/(1 - b), with the Cb/C2
  hole at n - W (psi updated, field untouched) as the reference has it
  (BaseFDTD11.py:364-393).
"""
import math

import numpy as np
import pytest

from refcases import product_objects, small_setup

pytestmark = pytest.mark.gpu

Z0 = 376.730313668


def _transverse_sheet(N, lo, hi, vx, npart, mass_scale=1.0, density=1e16):
    from fdtd_pic_amd import configs
    rng = np.random.default_rng(21)
    x = lo + rng.uniform(0, hi - lo, npart)
    beta = np.zeros((npart, 3))
    beta[:, 0] = vx / configs.C0
    u = beta / np.sqrt(1 - np.sum(beta ** 2, axis=1))[:, None]
    L = (hi - lo) * configs.DZ
    w = density * L / npart
    return x, u, -configs.E_CHARGE, configs.M_ELECTRON * mass_scale, w, L


@pytest.mark.parametrize("prec,tol", [("f64", 1e-12), ("f32", 1e-5)])
def test_device_deposit_equals_surface_current(prec, tol):
    from fdtd_pic_amd import configs, engine
    p = configs.base(600, 1, slab_start=300, slab_stop=300, src_amp=0.0)
    n, vx = 1e16, 0.3 * configs.C0
    x, u, q, m, w, L = _transverse_sheet(600, 300, 340, vx, 4000, density=n)
    with engine.Context(prec) as ctx:
        sim = ctx.add(engine.physical_spec(p), beam=engine.Beam(x, u, q, m, w, configs.DT, configs.DZ))
        ctx.run_all()
        J = sim.read("J")
    expect = Z0 * n * q * vx * L
    assert abs(J.sum() - expect) < tol * abs(expect)


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_device_current_sheet_radiates_minus_half_J(prec):
    from fdtd_pic_amd import configs, engine
    N, n0, steps = 2000, 1000, 1700
    p = configs.base(N, steps, slab_start=n0, slab_stop=n0, src_amp=0.0, probes=[n0 - 400, n0 + 400])
    n, vx = 1e16, 0.1 * configs.C0
    x, u, q, m, w, L = _transverse_sheet(N, n0 - 2, n0 + 2, vx, 2000, mass_scale=1e12, density=n)
    with engine.Context(prec) as ctx:
        sim = ctx.add(engine.physical_spec(p), beam=engine.Beam(x, u, q, m, w, configs.DT, configs.DZ))
        ctx.run_all()
        E, H = sim.read("E"), sim.read("H")
    half = -Z0 * n * q * vx * L / 2            # E = -J'/2
    assert abs(E[n0 - 400] / half - 1) < 1e-2 and abs(E[n0 + 400] / half - 1) < 1e-2
    assert abs(H[n0 - 400] / half - 1) < 1e-2 and abs(H[n0 + 399] / half + 1) < 1e-2


def _product_cpml():
    import BaseFDTD11
    import Solver_Engine
    setup = small_setup(Nz=300, T=10, W=20, F=200, R=299)
    V, P, C_V, C_P = product_objects(setup, 9e9, LorentzMed=True)
    BaseFDTD11.FieldInit(V, P)
    BaseFDTD11.EmptySpaceCalc(V, P)
    BaseFDTD11.CPML_FieldInit(V, P, C_V, C_P)
    Solver_Engine.boundCondManager(V, P, C_V, C_P)
    return V, P, C_V, C_P


def test_device_psi_e_recursion_geometric_sum():
    import BaseFDTD11
    V, P, C_V, C_P = _product_cpml()
    n, W = P.Nz + 1, P.pmlWidth
    rng = np.random.default_rng(5)
    V.Hy = rng.normal(0, 1, n)
    dH = np.r_[0.0, V.Hy[1:] - V.Hy[:-1]]
    Ex0 = rng.normal(0, 1, n)
    V.Ex = Ex0.copy()
    C_V.psi_Ex = np.zeros(n)
    steps = 7
    for _ in range(steps):
        BaseFDTD11.CPML_Psi_e_Update(V, P, C_V, C_P)
    b, c = C_V.beX, C_V.ceX
    cells = np.r_[1:W, n - W:n]
    geo = (1 - b[cells] ** steps) / (1 - b[cells])
    cum = (steps - b[cells] * geo) / (1 - b[cells])
    np.testing.assert_allclose(C_V.psi_Ex[cells], c[cells] * dH[cells] * geo, rtol=1e-12)
    np.testing.assert_allclose(V.Ex[cells], Ex0[cells] - C_V.Cb[cells] * c[cells] * dH[cells] * cum, rtol=1e-12)
    hole = n - W
    assert C_V.psi_Ex[hole] != 0 and V.Ex[hole] == Ex0[hole]
    rest = np.setdiff1d(np.arange(n), cells)
    assert np.all(C_V.psi_Ex[rest] == 0) and np.array_equal(V.Ex[rest], Ex0[rest])


def test_device_psi_m_recursion_geometric_sum():
    import BaseFDTD11
    V, P, C_V, C_P = _product_cpml()
    n, W = P.Nz + 1, P.pmlWidth
    rng = np.random.default_rng(6)
    V.Ex = rng.normal(0, 1, n)
    dE = np.r_[V.Ex[1:] - V.Ex[:-1], 0.0]
    H0 = rng.normal(0, 1, n)
    V.Hy = H0.copy()
    C_V.psi_Hy = np.zeros(n)
    steps = 5
    for _ in range(steps):
        BaseFDTD11.CPML_Psi_m_Update(V, P, C_V, C_P)
    b, c = C_V.bmY, C_V.cmY
    cells = np.r_[1:W, n - W:n - 1]
    geo = (1 - b[cells] ** steps) / (1 - b[cells])
    cum = (steps - b[cells] * geo) / (1 - b[cells])
    np.testing.assert_allclose(C_V.psi_Hy[cells], c[cells] * dE[cells] * geo, rtol=1e-12)
    np.testing.assert_allclose(V.Hy[cells], H0[cells] + C_V.C2[cells] * c[cells] * dE[cells] * cum, rtol=1e-12)
    assert C_V.psi_Hy[n - 1] == 0 and V.Hy[n - 1] == H0[n - 1]
    assert math.isfinite(float(np.sum(V.Hy)))
