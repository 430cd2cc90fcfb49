"""Drop-in for the reference's BaseFDTD.py, with that module's own semantics.

BaseFDTD.py is the pre-numba ancestor of BaseFDTD11.py.  Its functions keep
BaseFDTD11's names but compute different things, and this module gives each
the older module's result:

* set-up on the host, as the reference computes it: FieldInit's guards
  (Nz <= 3000, timeSteps <= 2**14) and 11-tuple (BaseFDTD.py:35-79);
  SmoothTurnOn's one-period sine that is cut at its second zero crossing
  (:82-106); vacuum coefficients that leave index Nz at zero (:109-118,
  :152-162); zero-filled CPML arrays (:176-197) and the older profile,
  reflection-coefficient, update-coefficient and denominator loops
  (:199-329), including their index mirrors and the alpha_Hy expression as
  written (:222);
* per-step updates as FP64 device kernels (fdtd_step_op, no FMA
  contraction, the reference's expression order and index ranges), so the
  results equal the reference's bit for bit: the E/H updates with the
  material self and coupling factors (:349-358, :392-408), the CPML psi
  recursions with the psi_E correction added and the psi_H difference taken
  as Ex[n] - Ex[n+1] (:331-344), the polarisation with its coefficients
  formed from gammaE, omega_0E, plasmaFreqE (:378-390), Dx and Ex
  (:411-422);
* the TFSF corrections, PEC/PMC ends and the polarisation history shift are
  single-cell or copy bookkeeping on the caller's arrays (:164-170,
  :361-375).

The five functions that are empty in the reference (ADE_MyUpdate,
ADE_NonLinMyUpdate, ADE_NonLinPxUpdate, SpatialFiltering,
SymbolicRegression) do nothing here either.  The per-step functions need
the HIP library (there is no host fallback).
"""
import sys

import numpy as np

GRID_LIMIT = 3000          # BaseFDTD.py:47
STEP_LIMIT = 2 ** 14       # BaseFDTD.py:44


def sourceGen(T):
    """BaseFDTD.py:27-29."""
    return np.exp(-(T - 30) * (T - 30) / 100)


def sourceGen2(P, T):
    """BaseFDTD.py:31-33."""
    return np.sin(2 * np.pi * P.freq_in * P.delT * T)


def FieldInit(V, P):
    """BaseFDTD.py:35-79.  A non-int Nz or timeSteps raises TypeError; the
    range guards print their message and exit, in the reference's order."""
    try:
        if type(P.Nz) != int:
            raise TypeError('Grid spaces must be a positive integer value')
        if type(P.timeSteps) != int:
            raise TypeError('Number of timeSteps must be positive integer valued')
        for bad, msg in ((P.timeSteps > STEP_LIMIT, 'timeSteps max too large'),
                         (P.Nz > GRID_LIMIT, 'Grid size too big'),
                         (P.Nz == 0, 'Cannot have grid size 0!'),
                         (P.timeSteps == 0, 'Cannot have zero timeSteps!'),
                         (P.Nz < 0, 'Grid size cannot be negative'),
                         (P.timeSteps < 0, 'Cannot have negative timeSteps')):
            if bad:
                raise ValueError(msg)
    except ValueError as e:
        print(e)
        sys.exit()
    n = P.Nz + 1
    V.Ex = np.zeros(n)
    V.Hy = np.zeros(n)
    # one shared empty list per history, as [[]] * timeSteps builds it
    V.Ex_History = [[]] * P.timeSteps
    V.Hy_History = [[]] * P.timeSteps
    V.Psi_Ex_History = [[]] * P.timeSteps
    V.Psi_Hy_History = [[]] * P.timeSteps
    V.Exs = []
    V.Hys = []
    V.polarisationCurr = np.zeros(n)
    V.Dx = np.zeros(n)
    V.tempVarPol = np.zeros(n)
    V.tempTempVarPol = np.zeros(n)
    return (V.tempTempVarPol, V.polarisationCurr, V.Ex, V.Dx, V.Ex_History, V.Hy, V.Hy_History,
            V.Psi_Ex_History, V.Psi_Hy_History, V.Exs, V.Hys)


def SmoothTurnOn(V, P):
    """BaseFDTD.py:82-106: Exs[t] = sin(2 pi / ppw * (Sc t)), Hys[t] the same
    at t + 1, from the first sample that turns positive after the sine went
    negative onwards both are zero (that sample's Hys too)."""
    steps = P.timeSteps
    V.Exs = np.zeros(steps)
    V.Hys = np.zeros(steps)
    ppw = P.c0 / (P.freq_in * P.dz)
    w = 2.0 * np.pi / ppw
    went_negative = cut = False
    p1 = p2 = 0
    for t in range(steps):
        V.Exs[t] = np.sin(w * (P.courantNo * t))
        V.Hys[t] = np.sin(w * (P.courantNo * (t + 1)))
        if V.Exs[t] < 0 and not went_negative:
            p1, went_negative = t, True
            print(p1, "p1Ind")
        if V.Exs[t] > 0 and went_negative and not cut:
            p2, cut = t, True
            print(p2, "p2Ind")
        if cut:
            V.Exs[t] = 0.0
            V.Hys[t] = 0.0
            V.Hys[p2] = 0.0
    return V.Exs, V.Hys


def _backgrounds(P):
    return (1 / P.CharImp) * P.courantNo, P.CharImp * P.courantNo


def EmptySpaceCalc(V, P):
    """BaseFDTD.py:109-118: the vacuum factors on cells 0 .. Nz-1; index Nz
    stays zero."""
    hb, eb = _backgrounds(P)
    V.UpHyMat = np.zeros(P.Nz + 1)
    V.UpExMat = np.zeros(P.Nz + 1)
    V.UpExMat[:P.Nz] = eb
    V.UpHyMat[:P.Nz] = hb
    return V.UpHyMat, V.UpExMat


def Material(V, P):
    """BaseFDTD.py:120-146: epsilon, mu and the four self/coupling factors per
    cell (vacuum outside [front, rear), the slab's values inside)."""
    V.epsilon = np.ones(P.Nz)
    V.mu = np.ones(P.Nz)
    front, rear = int(P.materialFrontEdge), P.materialRearEdge
    for k in range(P.Nz):
        inside = front <= k < int(rear)
        if k < front or k >= rear or inside:
            V.epsilon[k] = P.epsRe if inside else 1
            V.mu[k] = P.muRe if inside else 1
            V.UpExHcompsCo[k] = P.eHcompsCo if inside else 1
            V.UpExSelf[k] = P.eSelfCo if inside else 1
            V.UpHyEcompsCo[k] = P.hEcompsCo if inside else 1
            V.UpHySelf[k] = P.hSelfCo if inside else 1
    return V.epsilon, V.mu, V.UpExHcompsCo, V.UpExSelf, V.UpHyEcompsCo, V.UpHySelf


def UpdateCoef(V, P):
    """BaseFDTD.py:152-162: background / epsilon and background / mu on cells
    0 .. Nz-1 (index Nz zero); returned, not stored."""
    hb, eb = _backgrounds(P)
    UpHyMat = np.zeros(P.Nz + 1)
    UpExMat = np.zeros(P.Nz + 1)
    UpExMat[:P.Nz] = eb / np.asarray(V.epsilon[:P.Nz], dtype=np.float64)
    UpHyMat[:P.Nz] = hb / np.asarray(V.mu[:P.Nz], dtype=np.float64)
    return UpHyMat, UpExMat


def HyTfSfCorr(V, P, counts):
    """BaseFDTD.py:164-166."""
    V.Hy[P.nzsrc - 1] -= V.Exs[counts] / P.CharImp
    return V.Hy[P.nzsrc - 1]


def ExTfSfCorr(V, P, counts):
    """BaseFDTD.py:168-170."""
    V.Ex[P.nzsrc] += V.Hys[counts]
    return V.Ex[P.nzsrc]


_CPML_ARRAYS = ("kappa_Ex", "kappa_Hy", "psi_Ex", "psi_Hy", "alpha_Ex", "alpha_Hy", "sigma_Ex", "sigma_Hy",
                "beX", "bmY", "ceX", "cmY", "eLoss_CPML", "mLoss_CPML", "Ca", "Cb", "Cc", "C1", "C2", "C3")


def CPML_FieldInit(V, P, C_V, C_P):
    """BaseFDTD.py:176-197: twenty zeroed arrays of Nz + 1 (kappa included)."""
    for name in _CPML_ARRAYS:
        setattr(C_V, name, np.zeros(P.Nz + 1))
    return C_V


def CPML_ScalingCalc(V, P, C_V, C_P):
    """BaseFDTD.py:199-240.  Left strip: polynomial grading from the outer
    cell; right strip (cells >= Nz + 2 - W): copies of the left values taken
    with a descending index that starts at W (E) and W - 1 (H) - the
    sequence reads cells already written earlier in this loop."""
    W = P.pmlWidth
    g = lambda a: (a / (W - 1)) ** C_P.r_scale
    je, jh = W, W - 1
    for nz in range(P.Nz):
        k = nz + 1
        if k <= W:
            C_V.kappa_Ex[nz] = 1 + (C_P.kappaMax - 1) * g(W - k)
            C_V.sigma_Ex[nz] = C_P.sigmaOpt * g(W - k)
            C_V.alpha_Ex[nz] = C_P.alphaMax * (k / (W - 1)) ** C_P.r_a_scale
        elif nz >= P.Nz + 2 - W:
            C_V.kappa_Ex[nz], C_V.sigma_Ex[nz], C_V.alpha_Ex[nz] = \
                C_V.kappa_Ex[je], C_V.sigma_Ex[je], C_V.alpha_Ex[je]
            je -= 1
        else:
            C_V.kappa_Ex[nz], C_V.sigma_Ex[nz], C_V.alpha_Ex[nz] = 1, 0, 0
        if k <= W - 1:
            C_V.kappa_Hy[nz] = 1 + (C_P.kappaMax - 1) * g(W - k - 0.5)
            C_V.sigma_Hy[nz] = C_P.sigmaOpt * g(W - k - 0.5)
            # as written: k - (0.5 / (W - 1)), not (k - 0.5) / (W - 1)
            C_V.alpha_Hy[nz] = C_P.alphaMax * (k - 0.5 / (W - 1)) ** C_P.r_a_scale
        elif nz >= P.Nz + 2 - W:
            C_V.kappa_Hy[nz], C_V.sigma_Hy[nz], C_V.alpha_Hy[nz] = \
                C_V.kappa_Hy[jh], C_V.sigma_Hy[jh], C_V.alpha_Hy[jh]
            jh -= 1
        else:
            C_V.kappa_Hy[nz], C_V.sigma_Hy[nz], C_V.alpha_Hy[nz] = 1, 0, 0
    return C_V.sigma_Ex, C_V.sigma_Hy, C_V.alpha_Ex, C_V.alpha_Hy, C_V.kappa_Ex, C_V.kappa_Hy


def _ce(b, sigma, alpha, kappa):
    return (b - 1) * sigma / (sigma + alpha * kappa) / kappa


def CPML_Ex_RC_Define(V, P, C_V, C_P):
    """BaseFDTD.py:243-267: beX = exp(-(sigma dt / kappa + alpha dt / eps0))
    on both strips, zero between; ceX from the cell's own values on the left
    (zero at cell W-1 when sigma and alpha vanish there) and from the
    descending mirror index W-1, W-2, ... on the right."""
    W, N = P.pmlWidth, P.Nz
    j = W - 1
    s, a, kap = C_V.sigma_Ex, C_V.alpha_Ex, C_V.kappa_Ex
    for nz in range(N):
        left, right = nz <= W - 1, nz >= N + 2 - W
        if left or right:
            C_V.beX[nz] = np.exp(-((s[nz] * P.delT / kap[nz]) + ((a[nz] * P.delT) / P.permit_0)))
        else:
            C_V.beX[nz] = 0
        if s[nz] == 0 and nz == W - 1 and a[nz] == 0:
            C_V.ceX[nz] = 0
        elif left:
            C_V.ceX[nz] = _ce(C_V.beX[nz], s[nz], a[nz], kap[nz])
        elif right:
            C_V.ceX[nz] = _ce(C_V.beX[j], s[j], a[j], kap[j])
            j -= 1
        else:
            C_V.ceX[nz] = 0
    return C_V.beX, C_V.ceX


def CPML_HY_RC_Define(V, P, C_V, C_P):
    """BaseFDTD.py:270-292: bmY = exp(-(sigma / (kappa + alpha)) dt / mu0) on
    cells <= W-2 and, mirrored from index W-2 down, on the right strip; cmY
    like ceX (left to cell W-1, right from the descending index W-2)."""
    W, N = P.pmlWidth, P.Nz
    jb = jc = W - 2
    s, a, kap = C_V.sigma_Hy, C_V.alpha_Hy, C_V.kappa_Hy
    bm = lambda i: np.exp(-(s[i] / (kap[i] + a[i])) * P.delT / P.permea_0)
    for nz in range(N):
        right = nz >= N + 2 - W
        if nz <= W - 2:
            C_V.bmY[nz] = bm(nz)
        elif right:
            C_V.bmY[nz] = bm(jb)
            jb -= 1
        else:
            C_V.bmY[nz] = 0
        if s[nz] == 0 and nz == W - 1 and a[nz] == 0:
            C_V.cmY[nz] = 0
        elif nz <= W - 1:
            C_V.cmY[nz] = _ce(C_V.bmY[nz], s[nz], a[nz], kap[nz])
        elif right:
            C_V.cmY[nz] = _ce(C_V.bmY[jc], s[jc], a[jc], kap[jc])
            jc -= 1
        else:
            C_V.cmY[nz] = 0
    return C_V.bmY, C_V.cmY


def CPML_Ex_Update_Coef(V, P, C_V, C_P):
    """BaseFDTD.py:294-301, cells 0 .. Nz-2."""
    m = P.Nz - 1
    C_V.eLoss_CPML[:m] = (C_V.sigma_Ex[:m] * P.delT) / (2 * P.permit_0)
    C_V.Ca[:m] = V.UpExSelf[:m]
    C_V.Cb[:m] = V.UpExHcompsCo[:m] * V.UpExMat[:m]
    C_V.Cc[:m] = P.delT / ((1 + C_V.eLoss_CPML[:m]) * P.permit_0)
    return C_V.eLoss_CPML, C_V.Ca, C_V.Cb, C_V.Cc


def CPML_Hy_Update_Coef(V, P, C_V, C_P):
    """BaseFDTD.py:304-309, cells 0 .. Nz-3 (mLoss_CPML stays as it is)."""
    m = P.Nz - 2
    C_V.C1[:m] = 1
    C_V.C2[:m] = -P.delT / P.permea_0
    C_V.C3[:m] = P.delT / ((1 + C_V.mLoss_CPML[:m]) * P.permea_0)
    return C_V.mLoss_CPML, C_V.C1, C_V.C2, C_V.C3


def denominators(V, P, C_V, C_P):
    """BaseFDTD.py:311-329: 1/kappa on the strips (the right strip from the
    descending mirror index), 1 between, into the caller's den_Hydz and
    den_Exdz."""
    W, N = P.pmlWidth, P.Nz
    for den, kap, edge in ((C_V.den_Hydz, C_V.kappa_Hy, W - 2), (C_V.den_Exdz, C_V.kappa_Ex, W - 1)):
        j = edge
        for n in range(N):
            if n <= edge:
                den[n] = 1 / (kap[n])
            elif n >= N + 2 - W:
                den[n] = 1 / (kap[j])
                j -= 1
            else:
                den[n] = 1 / (1)
    return C_V.den_Exdz, C_V.den_Hydz


# ---- per-step updates: FP64 device kernels --------------------------------

def _run(op, arrays, scalars=(), ranges=((0, 0), (0, 0))):
    from fdtd_pic_amd import single_step
    return single_step.run_op(op, arrays, scalars, ranges)


def _put(obj, name, buf):
    from fdtd_pic_amd import single_step
    return single_step._store(obj, name, buf)


def _span(lo, hi):
    lo, hi = int(lo), int(hi)
    if lo < 0 and hi > lo:
        raise ValueError(f"index range [{lo}, {hi}) starts below 0 (the reference wraps to the array's end)")
    return lo, hi


def CPML_Psi_e_Update(V, P, C_V, C_P):
    """BaseFDTD.py:331-336, cells 1 .. Nz-2."""
    b = _run(10, [V.Ex, C_V.psi_Ex, V.Hy, C_V.beX, C_V.ceX, C_V.Cb], (), ((1, P.Nz - 1), (0, 0)))
    return _put(C_V, "psi_Ex", b[1]), _put(V, "Ex", b[0])


def CPML_Psi_m_Update(V, P, C_V, C_P):
    """BaseFDTD.py:339-344, cells 0 .. Nz-3."""
    b = _run(11, [V.Hy, C_V.psi_Hy, V.Ex, C_V.bmY, C_V.cmY, C_V.C2], (), ((0, P.Nz - 2), (0, 0)))
    return _put(C_V, "psi_Hy", b[1]), _put(V, "Hy", b[0])


def CPML_HyUpdate(V, P, C_V, C_P):
    """BaseFDTD.py:349-352, cells 0 .. Nz-3."""
    b = _run(1, [V.Hy, V.Ex, V.UpHySelf, V.UpHyMat, V.UpHyEcompsCo], (), ((0, P.Nz - 2), (0, 0)))
    return _put(V, "Hy", b[0])


def CPML_ExUpdate(V, P, C_V, C_P):
    """BaseFDTD.py:355-358, cells 1 .. Nz-2."""
    b = _run(9, [V.Ex, V.Hy, V.UpExSelf, V.UpExHcompsCo, V.UpExMat], (), ((1, P.Nz - 1), (0, 0)))
    return _put(V, "Ex", b[0])


def CPML_PEC(V, P, C_V, C_P):
    """BaseFDTD.py:361-363."""
    V.Ex[0] = 0
    return V.Ex[0], V.Ex[P.Nz - 1]


def CPML_PMC(V, P, C_V, C_P):
    """BaseFDTD.py:365-367."""
    V.Hy[P.Nz - 1] = 0
    return V.Hy[P.Nz - 1]


def ADE_TempPolCurr(V, P):
    """BaseFDTD.py:369-373: the history shift on cells 1 .. Nz-2 only."""
    s = slice(1, P.Nz - 1)
    V.tempTempVarPol[s] = V.tempVarPol[s]
    V.tempVarPol[s] = V.polarisationCurr[s]
    return V.tempTempVarPol, V.tempVarPol


def ade_coefficients(V, P):
    """A, B, C of BaseFDTD.py:379-387 (denominator 1/dt^2 + gamma/(2 dt))."""
    dt = P.delT
    D = (1 / dt ** 2) + (V.gammaE / (2 * dt))
    A = ((2 / dt ** 2) - V.omega_0E ** 2) / D
    B = ((V.gammaE / (2 * dt)) - 1 / dt ** 2) / D
    C = (P.permit_0 * (V.plasmaFreqE ** 2)) / D
    return A, B, C


def ADE_PolarisationCurrent_Ex(V, P, C_V, C_P):
    """BaseFDTD.py:378-390, cells front-1 .. rear-1."""
    A, B, C = ade_coefficients(V, P)
    rg = _span(P.materialFrontEdge - 1, P.materialRearEdge)
    b = _run(12, [V.polarisationCurr, V.tempVarPol, V.tempTempVarPol, V.Ex], (A, B, C), (rg, (0, 0)))
    return _put(V, "polarisationCurr", b[0])


def ADE_HyUpdate(V, P, C_V, C_P):
    """BaseFDTD.py:392-395, cells 1 .. Nz-2 (no coupling factor: times 1.0)."""
    one = np.ones(len(V.Hy))
    b = _run(1, [V.Hy, V.Ex, V.UpHySelf, V.UpHyMat, one], (), ((1, P.Nz - 1), (0, 0)))
    return _put(V, "Hy", b[0])


def ADE_MyUpdate():
    """Empty in the reference (BaseFDTD.py:397-399)."""


def ADE_ExUpdate(V, P, C_V, C_P):
    """BaseFDTD.py:401-408: cells 1 .. front-2, and rear-1 .. Nz-1 when
    rear < Nz - 1 (the slab and the cell before it are skipped)."""
    first = (1, int(P.materialFrontEdge - 1))
    second = _span(P.materialRearEdge - 1, P.Nz) if P.materialRearEdge < P.Nz - 1 else (0, 0)
    if second[0] == 0 and second[1] > 0:
        raise ValueError("rear edge 1 reads Hy[-1] (wrap-around) in the reference")
    one = np.ones(len(V.Ex))
    b = _run(9, [V.Ex, V.Hy, V.UpExSelf, V.UpExMat, one], (), (first, second))
    return _put(V, "Ex", b[0])


def ADE_ExCreate(V, P, C_V, C_P):
    """BaseFDTD.py:411-414, cells front-1 .. rear-1."""
    rg = _span(P.materialFrontEdge - 1, P.materialRearEdge)
    b = _run(5, [V.Ex, V.Dx, V.polarisationCurr], (P.permit_0,), (rg, (0, 0)))
    return _put(V, "Ex", b[0])


def ADE_DxUpdate(V, P, C_V, C_P):
    """BaseFDTD.py:417-422: Dx += dH dt / dz, then Ex = (Dx - P) / eps0, cells
    front-1 .. rear-1; returns (Dx, Ex).  A range from cell 0 reads Hy[-1],
    Python's last element: that cell is done here on the host the same way."""
    lo, hi = _span(P.materialFrontEdge - 1, P.materialRearEdge)
    Dx = np.ascontiguousarray(V.Dx, dtype=np.float64)
    Ex = np.ascontiguousarray(V.Ex, dtype=np.float64)
    Hy = np.asarray(V.Hy, dtype=np.float64)
    if lo == 0 and hi > 0:
        Dx[0] = Dx[0] + (Hy[0] - Hy[-1]) * P.delT / P.dz
        Ex[0] = (Dx[0] - V.polarisationCurr[0]) / P.permit_0
        lo = 1
    b = _run(13, [Dx, Ex, Hy, V.polarisationCurr], (P.delT, P.dz, P.permit_0), ((lo, hi), (0, 0)))
    return _put(V, "Dx", b[0]), _put(V, "Ex", b[1])


def ADE_NonLinMyUpdate():
    """Empty in the reference (BaseFDTD.py:425-426)."""


def ADE_NonLinPxUpdate():
    """Empty in the reference (BaseFDTD.py:428-429)."""


def SpatialFiltering():
    """Empty in the reference (BaseFDTD.py:431-433)."""


def SymbolicRegression():
    """Empty in the reference (BaseFDTD.py:435-436)."""
