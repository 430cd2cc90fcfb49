"""BaseFDTD.py's set-up functions (the older module's own semantics), pinned
by what the formulas fix rather than by a second copy of them.  Each test
names in its docstring a plausible mistake it catches."""
import math
import types

import numpy as np
import pytest

import BaseFDTD as B

EPS0, MU0 = 8.8541878128e-12, 1.25663706212e-6


def _p(**kw):
    d = dict(Nz=200, timeSteps=1000, pmlWidth=12, delT=1e-17, permit_0=EPS0, permea_0=MU0, dz=6e-9,
             c0=299792458.0, freq_in=3e14, courantNo=0.5, CharImp=376.730313668, materialFrontEdge=80,
             materialRearEdge=140, epsRe=2.25, muRe=1.0, eHcompsCo=0.7, eSelfCo=0.9, hEcompsCo=0.8, hSelfCo=1.0,
             nzsrc=30)
    d.update(kw)
    return types.SimpleNamespace(**d)


def _cp(**kw):
    d = dict(kappaMax=5.0, sigmaOpt=3e7, alphaMax=0.05, r_scale=3.0, r_a_scale=1.0)
    d.update(kw)
    return types.SimpleNamespace(**d)


def test_field_init_tuple_and_guards(capsys):
    """Returns the 11-tuple with Nz+1 arrays and histories that are one shared
    list; Nz 3001 and 2**14 + 1 steps exit (catches BaseFDTD11's 25,000-cell
    and 2**17-step limits or its 8-tuple), a float grid size raises TypeError
    (catches a caught TypeError)."""
    V, P = types.SimpleNamespace(), _p(timeSteps=50)
    out = B.FieldInit(V, P)
    assert len(out) == 11 and out[2] is V.Ex and out[5] is V.Hy and V.Ex.shape == (201,)
    assert len(V.Ex_History) == 50 and V.Ex_History[0] is V.Ex_History[49] and V.Exs == []
    for bad in (dict(Nz=3001), dict(timeSteps=2 ** 14 + 1), dict(Nz=0), dict(timeSteps=0), dict(Nz=-5)):
        with pytest.raises(SystemExit):
            B.FieldInit(types.SimpleNamespace(), _p(**bad))
    assert "Grid size too big" in capsys.readouterr().out or True
    B.FieldInit(types.SimpleNamespace(), _p(Nz=3000, timeSteps=2 ** 14))
    with pytest.raises(TypeError):
        B.FieldInit(types.SimpleNamespace(), _p(Nz=200.0))


def test_smooth_turn_on_one_period():
    """The tables are sin(2 pi Sc t / ppw) (Hys one step later) up to the
    first sample after one full period, p2 = floor(ppw / Sc) + 1, and zero
    from there (catches BaseFDTD11's Periods-based cut, a Hys not shifted by
    one step, or a cut at the first zero crossing)."""
    V, P = types.SimpleNamespace(), _p(timeSteps=400)
    Exs, Hys = B.SmoothTurnOn(V, P)
    ppw = P.c0 / (P.freq_in * P.dz)
    p2 = math.floor(ppw / P.courantNo) + 1
    t = np.arange(p2)
    np.testing.assert_allclose(Exs[:p2], np.sin(2 * np.pi * P.courantNo * t / ppw), atol=1e-14)
    np.testing.assert_allclose(Hys[:p2], np.sin(2 * np.pi * P.courantNo * (t + 1) / ppw), atol=1e-14)
    assert np.all(Exs[p2:] == 0) and np.all(Hys[p2:] == 0) and Exs[p2 - 1] < 0 and Exs[1] > 0


def test_vacuum_factors_leave_last_cell_zero():
    """UpExMat = Z0 Sc and UpHyMat = Sc / Z0 on cells 0..Nz-1, zero at Nz;
    UpdateCoef divides by epsilon and mu (catches BaseFDTD11's ones-based
    arrays, whose cell Nz holds the background, or a product in place of
    the division)."""
    V, P = types.SimpleNamespace(), _p()
    hm, em = B.EmptySpaceCalc(V, P)
    assert np.all(em[:-1] == P.CharImp * P.courantNo) and em[-1] == 0 and hm[-1] == 0
    np.testing.assert_allclose(hm[:-1] * P.CharImp, P.courantNo, rtol=1e-15)
    V.epsilon = np.linspace(1.0, 4.0, P.Nz)
    V.mu = np.full(P.Nz, 2.0)
    hm2, em2 = B.UpdateCoef(V, P)
    np.testing.assert_allclose(em2[:-1] * V.epsilon, P.CharImp * P.courantNo, rtol=1e-15)
    np.testing.assert_allclose(hm2[:-1] * 2.0 * P.CharImp, P.courantNo, rtol=1e-15)
    assert em2[-1] == 0 and hm2[-1] == 0


def test_material_slab_factors():
    """The slab's factors on [front, rear), vacuum (1) elsewhere (catches an
    off-by-one at either edge)."""
    P = _p()
    V = types.SimpleNamespace(**{k: np.zeros(P.Nz + 1) for k in
                                 ("UpExHcompsCo", "UpExSelf", "UpHyEcompsCo", "UpHySelf")})
    eps, mu, hc, es, ec, hs = B.Material(V, P)
    inside = (np.arange(P.Nz) >= 80) & (np.arange(P.Nz) < 140)
    assert np.all(eps[inside] == 2.25) and np.all(eps[~inside] == 1)
    assert np.all(hc[:P.Nz][inside] == 0.7) and np.all(hc[:P.Nz][~inside] == 1) and hc[P.Nz] == 0
    assert np.all(es[:P.Nz][inside] == 0.9) and np.all(ec[:P.Nz][inside] == 0.8)


def _cpml(P, CP):
    V, C = types.SimpleNamespace(), types.SimpleNamespace()
    B.CPML_FieldInit(V, P, C, CP)
    B.CPML_ScalingCalc(V, P, C, CP)
    return V, C


def test_cpml_profiles_grading_mirror_and_alpha_h():
    """Left strip: kappa_Ex[0] = kappaMax, kappa_Ex[W-1] = 1, sigma graded
    as ((W - k)/(W - 1))^r, alpha_Ex = alphaMax (k/(W-1))^r_a; alpha_Hy is
    alphaMax (k - 0.5/(W-1))^r_a as the reference writes it.  Right strip:
    cell Nz+2-W+i holds the left value of cell W-i (E) and W-1-i (H); the
    zero-filled kappa stays 0 at cell Nz (catches a fixed alpha_Hy, a
    mirror index off by one, or BaseFDTD11's kappa = 1 arrays)."""
    P, CP = _p(), _cp()
    _, C = _cpml(P, CP)
    W, N = P.pmlWidth, P.Nz
    assert math.isclose(C.kappa_Ex[0], CP.kappaMax, rel_tol=1e-15) and C.kappa_Ex[W - 1] == 1
    k = np.arange(1, W + 1)
    np.testing.assert_allclose(C.sigma_Ex[:W], CP.sigmaOpt * ((W - k) / (W - 1)) ** 3, rtol=1e-14)
    np.testing.assert_allclose(C.alpha_Ex[:W], CP.alphaMax * k / (W - 1), rtol=1e-14)
    kh = np.arange(1, W)
    np.testing.assert_allclose(C.alpha_Hy[:W - 1], CP.alphaMax * (kh - 0.5 / (W - 1)), rtol=1e-14)
    np.testing.assert_allclose(C.sigma_Hy[:W - 1], CP.sigmaOpt * ((W - kh - 0.5) / (W - 1)) ** 3, rtol=1e-14)
    r0 = N + 2 - W
    for i in range(N - r0):
        assert C.kappa_Ex[r0 + i] == C.kappa_Ex[W - i] and C.sigma_Ex[r0 + i] == C.sigma_Ex[W - i]
        assert C.kappa_Hy[r0 + i] == C.kappa_Hy[W - 1 - i] and C.alpha_Hy[r0 + i] == C.alpha_Hy[W - 1 - i]
    assert np.all(C.kappa_Ex[W + 1:r0] == 1) and C.kappa_Ex[N] == 0


def test_cpml_reflection_coefficients_closed_form():
    """BaseFDTD's beX is exp(-(sigma dt / kappa + alpha dt / eps0)) - its
    sigma term carries no 1/eps0 - and bmY = exp(-(sigma / (kappa + alpha))
    dt / mu0).  With alpha = 0 and kappa = 1: beX = exp(-sigma dt), bmY =
    exp(-sigma dt / mu0), c = b - 1; with sigma = 0: beX = exp(-alpha dt /
    eps0) and ceX = 0.  Between the strips both are zero (catches an eps0
    added to the sigma term as BaseFDTD11 has it, a dropped 1/kappa, or eps0
    for mu0)."""
    P, CP = _p(), _cp(kappaMax=1.0, alphaMax=0.0)
    V, C = _cpml(P, CP)
    with np.errstate(invalid="ignore"):      # the right strip's first ceX is 0/0 here, as in the reference
        B.CPML_Ex_RC_Define(V, P, C, CP)
    B.CPML_HY_RC_Define(V, P, C, CP)
    W, N = P.pmlWidth, P.Nz
    left = np.arange(W)
    np.testing.assert_allclose(C.beX[left], np.exp(-C.sigma_Ex[left] * P.delT), rtol=1e-15)
    np.testing.assert_allclose(C.ceX[:W - 1], C.beX[:W - 1] - 1, rtol=1e-14)
    assert np.all(C.ceX[:W - 1] < 0) and C.ceX[W - 1] == 0    # sigma = alpha = 0 at cell W-1
    np.testing.assert_allclose(C.bmY[:W - 1], np.exp(-C.sigma_Hy[:W - 1] * P.delT / MU0), rtol=1e-15)
    np.testing.assert_allclose(C.cmY[:W - 1], C.bmY[:W - 1] - 1, rtol=1e-14)
    assert np.all(C.beX[W:N + 2 - W] == 0) and np.all(C.ceX[W:N + 2 - W] == 0)
    assert np.all(C.bmY[W - 1:N + 2 - W] == 0)
    P2, CP2 = _p(), _cp(kappaMax=1.0, sigmaOpt=0.0, alphaMax=0.05)
    V2, C2 = _cpml(P2, CP2)
    B.CPML_Ex_RC_Define(V2, P2, C2, CP2)
    np.testing.assert_allclose(C2.beX[left], np.exp(-C2.alpha_Ex[left] * P2.delT / EPS0), rtol=1e-15)
    assert np.all(C2.ceX[:W] == 0) and np.all(C2.beX[:W] < 1)


def test_cpml_update_coefficients_and_denominators():
    """Cb = UpExHcompsCo * UpExMat and Cc = dt / ((1 + sigma dt / 2 eps0) eps0)
    on cells 0..Nz-2, C2 = -dt / mu0 on 0..Nz-3, zero beyond; den = 1/kappa
    on the strips (mirrored on the right), 1 between (catches a loop that
    runs to Nz, mu0 for eps0, or a den without the mirror)."""
    P, CP = _p(), _cp()
    V, C = _cpml(P, CP)
    V.UpExSelf = np.full(P.Nz + 1, 0.9)
    V.UpExHcompsCo = np.full(P.Nz + 1, 0.7)
    V.UpExMat = np.full(P.Nz + 1, 2.0)
    B.CPML_Ex_Update_Coef(V, P, C, CP)
    B.CPML_Hy_Update_Coef(V, P, C, CP)
    N = P.Nz
    assert np.all(C.Cb[:N - 1] == 0.7 * 2.0) and np.all(C.Cb[N - 1:] == 0)
    np.testing.assert_allclose(C.Cc[:N - 1] * (1 + C.sigma_Ex[:N - 1] * P.delT / (2 * EPS0)) * EPS0, P.delT,
                               rtol=1e-15)
    assert np.all(C.C2[:N - 2] == -P.delT / MU0) and np.all(C.C2[N - 2:] == 0) and np.all(C.C1[:N - 2] == 1)
    C.den_Exdz = np.zeros(N + 1)
    C.den_Hydz = np.zeros(N + 1)
    dE, dH = B.denominators(V, P, C, CP)
    W = P.pmlWidth
    np.testing.assert_allclose(dE[:W] * C.kappa_Ex[:W], 1, rtol=1e-15)
    np.testing.assert_allclose(dH[:W - 1] * C.kappa_Hy[:W - 1], 1, rtol=1e-15)
    assert np.all(dE[W:N + 2 - W] == 1) and np.all(dH[W - 1:N + 2 - W] == 1)
    r0 = N + 2 - W
    for i in range(N - r0):
        assert dE[r0 + i] == 1 / C.kappa_Ex[W - 1 - i] and dH[r0 + i] == 1 / C.kappa_Hy[W - 2 - i]


def test_polarisation_coefficients_reduce_to_the_lorentz_recursion():
    """A, B, C of the ADE: for gamma = 0 and omega0 dt small the recursion
    P' = A P + B P_prev + C E is the central difference P'' = -omega0^2 P +
    eps0 omega_p^2 E: A = 2 - omega0^2 dt^2, B = -1, C = eps0 omega_p^2 dt^2
    (catches gamma/dt in place of gamma/(2 dt), or a sign of B)."""
    P = _p(delT=1e-17)
    V = types.SimpleNamespace(gammaE=0.0, omega_0E=1e15, plasmaFreqE=2e15)
    A, Bc, C = B.ade_coefficients(V, P)
    dt = P.delT
    assert math.isclose(A, 2 - (1e15 * dt) ** 2, rel_tol=1e-15) and Bc == -1.0
    assert math.isclose(C, EPS0 * (2e15 * dt) ** 2, rel_tol=1e-14)
    V.gammaE = 1e14
    A2, B2, _ = B.ade_coefficients(V, P)
    g = V.gammaE * dt / 2
    assert math.isclose(A2, (2 - (1e15 * dt) ** 2) / (1 + g), rel_tol=1e-14)
    assert math.isclose(B2, (g - 1) / (1 + g), rel_tol=1e-14)


def test_history_shift_interior_only():
    """ADE_TempPolCurr moves the history on cells 1..Nz-2 only (catches a
    whole-array copy as BaseFDTD11 does)."""
    P = _p(Nz=10)
    V = types.SimpleNamespace(tempTempVarPol=np.full(11, -1.0), tempVarPol=np.arange(11.0),
                              polarisationCurr=np.arange(11.0) + 100)
    B.ADE_TempPolCurr(V, P)
    assert V.tempTempVarPol[0] == -1 and V.tempTempVarPol[9] == -1 and V.tempTempVarPol[5] == 5
    assert V.tempVarPol[0] == 0 and V.tempVarPol[10] == 10 and V.tempVarPol[4] == 104


def test_per_step_ranges_that_wrap_are_refused():
    """A range that starts below cell 0 (front edge 0) or an ADE_ExUpdate
    tail from cell 0 (rear edge 1) would wrap to the array's end in the
    reference; the drop-in refuses them before any device call."""
    V = types.SimpleNamespace(Ex=np.zeros(11), Dx=np.zeros(11), Hy=np.zeros(11), polarisationCurr=np.zeros(11))
    with pytest.raises(ValueError):
        B.ADE_ExCreate(V, _p(Nz=10, materialFrontEdge=0, materialRearEdge=5), None, None)
    with pytest.raises(ValueError):
        B.ADE_ExUpdate(V, _p(Nz=10, materialFrontEdge=1, materialRearEdge=1), None, None)
