"""BaseFDTD.py's per-step functions (the older module's semantics), each run
as an FP64 device kernel by the drop-in, against the oracle's loop
restatement (oracle/faithful_base.py) on random state: bitwise equal.  Then
closed forms that a plausible mistake fails: the psi_E series under a frozen
difference (sign of the correction), the reversed psi_H difference, the
slab cells ADE_ExUpdate skips, the Hy[-1] read of a Dx range from cell 0,
and a 40-step CPML run composed of the set-up and per-step functions."""
import types

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NZ, FRONT, REAR = 300, 150, 260


def _state(seed, front=FRONT, rear=REAR):
    rng = np.random.default_rng(seed)
    n = NZ + 1
    P = types.SimpleNamespace(Nz=NZ, delT=1.7e-17, dz=1e-8, permit_0=8.8541878128e-12,
                              materialFrontEdge=front, materialRearEdge=rear)
    V = types.SimpleNamespace(gammaE=3e13, omega_0E=1.2e15, plasmaFreqE=2.1e15)
    for k in ("Ex", "Hy"):
        setattr(V, k, rng.normal(size=n))
    for k in ("Dx", "polarisationCurr", "tempVarPol", "tempTempVarPol"):
        setattr(V, k, rng.normal(size=n) * 1e-9)
    for k in ("UpExSelf", "UpExHcompsCo", "UpExMat", "UpHySelf", "UpHyMat", "UpHyEcompsCo"):
        setattr(V, k, rng.uniform(0.5, 2.0, n))
    C = types.SimpleNamespace(**{k: rng.normal(size=n) for k in
                                 ("psi_Ex", "psi_Hy", "beX", "ceX", "Cb", "bmY", "cmY", "C2")})
    return V, P, C


def _copy(ns):
    return types.SimpleNamespace(**{k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in vars(ns).items()})


def _same(a, b, tag):
    assert np.array_equal(np.asarray(a), np.asarray(b)), (tag, float(np.max(np.abs(np.asarray(a) - b))))


def _oracle(name, V, P, C):
    from oracle import faithful_base as O
    if name == "CPML_Psi_e_Update":
        return O.cpml_psi_e_update(V.Ex, C.psi_Ex, V.Hy, C.beX, C.ceX, C.Cb, NZ)
    if name == "CPML_Psi_m_Update":
        return O.cpml_psi_m_update(V.Hy, C.psi_Hy, V.Ex, C.bmY, C.cmY, C.C2, NZ)
    if name == "CPML_HyUpdate":
        return O.cpml_hy_update(V.Hy, V.Ex, V.UpHySelf, V.UpHyMat, V.UpHyEcompsCo, NZ)
    if name == "CPML_ExUpdate":
        return O.cpml_ex_update(V.Ex, V.Hy, V.UpExSelf, V.UpExHcompsCo, V.UpExMat, NZ)
    if name == "ADE_HyUpdate":
        return O.ade_hy_update(V.Hy, V.Ex, V.UpHySelf, V.UpHyMat, NZ)
    if name == "ADE_ExUpdate":
        return O.ade_ex_update(V.Ex, V.Hy, V.UpExSelf, V.UpExMat, P.materialFrontEdge, P.materialRearEdge, NZ)
    if name == "ADE_ExCreate":
        return O.ade_ex_create(V.Ex, V.Dx, V.polarisationCurr, P.permit_0, P.materialFrontEdge,
                               P.materialRearEdge)
    if name == "ADE_DxUpdate":
        return O.ade_dx_update(V.Dx, V.Ex, V.Hy, V.polarisationCurr, P.delT, P.dz, P.permit_0,
                               P.materialFrontEdge, P.materialRearEdge)
    if name == "ADE_PolarisationCurrent_Ex":
        return O.ade_polarisation_current_ex(V.polarisationCurr, V.tempVarPol, V.tempTempVarPol, V.Ex, V.gammaE,
                                             V.omega_0E, V.plasmaFreqE, P.delT, P.permit_0,
                                             P.materialFrontEdge, P.materialRearEdge)
    raise KeyError(name)


NAMES = ["CPML_Psi_e_Update", "CPML_Psi_m_Update", "CPML_HyUpdate", "CPML_ExUpdate", "ADE_HyUpdate",
         "ADE_ExUpdate", "ADE_ExCreate", "ADE_DxUpdate", "ADE_PolarisationCurrent_Ex"]


@pytest.mark.parametrize("edges", [(FRONT, REAR), (1, 40), (5, NZ - 1)])
@pytest.mark.parametrize("name", NAMES)
def test_base_op_bitwise(name, edges):
    import BaseFDTD
    print("BaseFDTD", name, edges, flush=True)
    V, P, C = _state(11, *edges)
    ref = _oracle(name, _copy(V), P, _copy(C))
    got = getattr(BaseFDTD, name)(V, P, C, None)
    ref = ref if isinstance(ref, tuple) else (ref,)
    got = got if isinstance(got, tuple) else (got,)
    assert len(ref) == len(got)
    for k, (r, g) in enumerate(zip(ref, got)):
        _same(g, r, (name, k))


def test_psi_e_frozen_difference_series():
    """psi_n = c dH (1 - b^n) / (1 - b) and Ex_n = Ex_0 + Cb sum_k psi_k on
    cells 1..Nz-2, cells 0 and Nz-1 untouched (catches BaseFDTD11's
    subtracted correction, or the range of the two-strip form)."""
    import BaseFDTD
    V, P, C = _state(3)
    b, c, cb = 0.7, 0.3, 0.5
    C.beX[:], C.ceX[:], C.Cb[:], C.psi_Ex[:] = b, c, cb, 0.0
    Ex0 = V.Ex.copy()
    dH = V.Hy[1:NZ - 1] - V.Hy[:NZ - 2]
    acc = np.zeros(NZ - 2)
    for n in range(1, 21):
        BaseFDTD.CPML_Psi_e_Update(V, P, C, None)
        psi = c * dH * (1 - b ** n) / (1 - b)
        acc += psi
        np.testing.assert_allclose(C.psi_Ex[1:NZ - 1], psi, rtol=1e-13, atol=1e-15)
    np.testing.assert_allclose(V.Ex[1:NZ - 1] - Ex0[1:NZ - 1], cb * acc, rtol=1e-12, atol=1e-13)
    assert V.Ex[0] == Ex0[0] and V.Ex[NZ - 1] == Ex0[NZ - 1]


def test_psi_m_reversed_difference():
    """From psi = 0 with Ex[n] = n, one update gives psi_Hy = cmY (Ex[n] -
    Ex[n+1]) = -cmY on cells 0..Nz-3 and Hy += C2 psi (catches the forward
    difference of BaseFDTD11)."""
    import BaseFDTD
    V, P, C = _state(4)
    V.Ex = np.arange(NZ + 1, dtype=np.float64)
    C.psi_Hy[:] = 0.0
    H0 = V.Hy.copy()
    BaseFDTD.CPML_Psi_m_Update(V, P, C, None)
    _same(C.psi_Hy[:NZ - 2], -C.cmY[:NZ - 2], "psi")
    _same(V.Hy[:NZ - 2], H0[:NZ - 2] + C.C2[:NZ - 2] * -C.cmY[:NZ - 2], "Hy")
    assert np.array_equal(V.Hy[NZ - 2:], H0[NZ - 2:])


def test_ex_update_skips_the_slab():
    """ADE_ExUpdate leaves cell 0 and cells front-1 .. rear-2 untouched and
    updates 1 .. front-2 and rear-1 .. Nz-1 (catches a loop over every cell,
    or rear - 1 taken as the first updated cell's neighbour)."""
    import BaseFDTD
    V, P, C = _state(5)
    E0 = V.Ex.copy()
    BaseFDTD.ADE_ExUpdate(V, P, C, None)
    changed = V.Ex != E0
    assert not changed[0] and not changed[FRONT - 1:REAR - 1].any()
    assert changed[1:FRONT - 1].all() and changed[REAR - 1:NZ].all() and not changed[NZ]


def test_dx_update_from_cell_zero_reads_the_last_hy():
    """A Dx range from cell 0 takes Hy[-1] (the last cell) as cell 0's left
    neighbour, as Python indexing does; Ex = (Dx - P) / eps0 on the range."""
    import BaseFDTD
    V, P, C = _state(6, 1, 30)
    D0 = V.Dx.copy()
    Dx, Ex = BaseFDTD.ADE_DxUpdate(V, P, C, None)
    assert Dx[0] == D0[0] + (V.Hy[0] - V.Hy[NZ]) * P.delT / P.dz
    _same(Ex[:30], (Dx[:30] - V.polarisationCurr[:30]) / P.permit_0, "Ex")


def test_composed_cpml_run_bitwise():
    """Forty steps of a vacuum grid with CPML ends built by BaseFDTD's set-up
    functions and advanced by its per-step functions on the device, against
    the same composition of the oracle's loops: bitwise, and the pulse
    entered the grid."""
    import BaseFDTD
    from oracle import faithful_base as O
    P = types.SimpleNamespace(Nz=NZ, timeSteps=40, pmlWidth=20, delT=1.6678e-17, permit_0=8.8541878128e-12,
                              permea_0=1.25663706212e-6, dz=1e-8, c0=299792458.0, freq_in=4e14, courantNo=0.5,
                              CharImp=376.730313668, materialFrontEdge=150, materialRearEdge=200, epsRe=1.0,
                              muRe=1.0, eHcompsCo=1.0, eSelfCo=1.0, hEcompsCo=1.0, hSelfCo=1.0, nzsrc=60)
    CP = types.SimpleNamespace(kappaMax=3.0, sigmaOpt=2e7, alphaMax=0.02, r_scale=3.0, r_a_scale=1.0)
    V = types.SimpleNamespace(**{k: np.zeros(NZ + 1) for k in
                                 ("UpExHcompsCo", "UpExSelf", "UpHyEcompsCo", "UpHySelf")})
    C = types.SimpleNamespace()
    BaseFDTD.FieldInit(V, P)
    BaseFDTD.EmptySpaceCalc(V, P)
    BaseFDTD.Material(V, P)
    for f in (BaseFDTD.CPML_FieldInit, BaseFDTD.CPML_ScalingCalc, BaseFDTD.CPML_Ex_RC_Define,
              BaseFDTD.CPML_HY_RC_Define, BaseFDTD.CPML_Ex_Update_Coef, BaseFDTD.CPML_Hy_Update_Coef):
        f(V, P, C, CP)
    BaseFDTD.SmoothTurnOn(V, P)
    o = {"Ex": V.Ex.copy(), "Hy": V.Hy.copy(), "pe": C.psi_Ex.copy(), "pm": C.psi_Hy.copy()}
    for t in range(P.timeSteps):
        BaseFDTD.CPML_HyUpdate(V, P, C, CP)
        BaseFDTD.CPML_Psi_m_Update(V, P, C, CP)
        BaseFDTD.CPML_ExUpdate(V, P, C, CP)
        BaseFDTD.CPML_Psi_e_Update(V, P, C, CP)
        V.Ex[P.nzsrc] += V.Exs[t]
        o["Hy"] = O.cpml_hy_update(o["Hy"], o["Ex"], V.UpHySelf, V.UpHyMat, V.UpHyEcompsCo, NZ)
        o["pm"], o["Hy"] = O.cpml_psi_m_update(o["Hy"], o["pm"], o["Ex"], C.bmY, C.cmY, C.C2, NZ)
        o["Ex"] = O.cpml_ex_update(o["Ex"], o["Hy"], V.UpExSelf, V.UpExHcompsCo, V.UpExMat, NZ)
        o["pe"], o["Ex"] = O.cpml_psi_e_update(o["Ex"], o["pe"], o["Hy"], C.beX, C.ceX, C.Cb, NZ)
        o["Ex"][P.nzsrc] += V.Exs[t]
    for k, a in (("Ex", V.Ex), ("Hy", V.Hy), ("pe", C.psi_Ex), ("pm", C.psi_Hy)):
        _same(a, o[k], k)
    assert np.count_nonzero(V.Ex) > 40 and np.abs(V.Hy).max() > 0
