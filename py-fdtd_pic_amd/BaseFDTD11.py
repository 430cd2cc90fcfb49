"""Drop-in for the reference's BaseFDTD11.py.

Set-up functions (field allocation, vacuum/material coefficients, CPML
profiles and coefficients, sources, analytic reflection) run on the host
exactly as the reference computes them (BaseFDTD11.py:36-357, :882-923),
because the device loop consumes the arrays they produce.

The per-step update kernels of the reference (ADE_ExUpdate, ADE_HyUpdate,
CPML_Psi_e_Update, CPML_Psi_m_Update, ADE_DxUpdate, ADE_ExCreate,
ADE_PolarisationCurrent_Ex, AcubicFinder/Nonlin_Cubic_Solver,
NonLinExUpdate) are fused into the GPU step loop that Solver_Engine's
integrators drive; called one at a time they run one fused device step
through `fdtd_pic_amd.single_step` and copy the arrays back (slow by
construction: the boundary of the fast path is the integrator).
"""
import sys

import numpy as np

# Reference grid guard (BaseFDTD11.py:45-49); raise it to run larger grids.
GRID_LIMIT = 25000
STEP_LIMIT = 2 ** 17


def sourceGen(T):
    return np.exp(-(T - 30) * (T - 30) / 100)


def sourceGen2(P, T):
    return np.sin(2 * np.pi * P.freq_in * P.delT * T)


def FieldInit(V, P):
    """BaseFDTD11.FieldInit (:36-82): guards, then zeroed state arrays."""
    try:
        if type(P.Nz) != int:
            raise TypeError('Grid spaces must be a positive integer value')
        if type(P.timeSteps) != int:
            raise TypeError('Number of timeSteps must be positive integer valued')
        if P.timeSteps > STEP_LIMIT:
            raise ValueError('timeSteps max too large')
        if P.Nz > GRID_LIMIT:
            raise ValueError('Grid size too big')
        if P.Nz == 0:
            raise ValueError('Cannot have grid size 0!')
        if P.timeSteps == 0:
            raise ValueError('Cannot have zero timeSteps!')
        if P.Nz < 0:
            raise ValueError('Grid size cannot be negative')
        if P.timeSteps < 0:
            raise ValueError('Cannot have negative timeSteps')
    except ValueError as e:
        print(e)
        sys.exit()
    n = P.Nz + 1
    V.Ex = np.zeros(n)
    V.Hy = np.zeros(n)
    V.polarisationCurr = np.zeros(n)
    V.Dx = np.zeros(n)
    V.tempVarPol = np.zeros(n)
    V.tempTempVarPol = np.zeros(n)
    V.tempVarE = np.zeros(n)
    V.tempTempVarE = np.zeros(n)
    return (V.tempVarPol, V.tempTempVarE, V.tempVarE, V.tempTempVarPol, V.polarisationCurr, V.Ex,
            V.Dx, V.Hy)


def Gaussian(V, P):
    """BaseFDTD11.Gaussian (:86-96), step-index time and cos inside the exponent as written."""
    t = np.arange(P.timeSteps)
    fc = 200
    tau = fc * 2.2
    gaussArg = ((t - tau) * (t - tau)) / (fc * fc) * np.cos(2 * np.pi * P.freq_in * (t - tau))
    out = np.exp(-gaussArg) * 2
    return P.Amplitude * out / P.courantNo


def SineCont(V, P, rep, interv):
    """Empty in the reference (BaseFDTD11.py:98-99)."""


def Ricker(V, P):
    """Empty in the reference (BaseFDTD11.py:101-102)."""


def SmoothTurnOn(V, P, tempfreq=0):
    """BaseFDTD11.SmoothTurnOn (:104-120): sine for Periods periods, then zero."""
    frq = tempfreq if tempfreq != 0 else P.freq_in
    ppw = P.c0 / (frq * P.dz)
    Exs, Hys = [], []
    limit = P.period * P.Periods
    for timer in range(P.timeSteps):
        if timer * P.delT < limit:
            Exs.append(float(np.sin(2.0 * np.pi / ppw * (P.courantNo * timer))))
            Hys.append(float(np.sin(2.0 * np.pi / ppw * (P.courantNo * (timer + 1)))))
        else:
            Exs.append(0)
            Hys.append(0)
    return Exs, Hys


def EmptySpaceCalc(V, P):
    """BaseFDTD11.EmptySpaceCalc (:124-132): UpExMat = Z0*Sc, UpHyMat = Sc/Z0."""
    V.UpHyMat = np.ones(P.Nz + 1)
    V.UpExMat = np.ones(P.Nz + 1)
    V.UpExMat *= P.CharImp * P.courantNo
    V.UpHyMat *= (1 / P.CharImp) * P.courantNo
    return V.UpHyMat, V.UpExMat


def Material(V, P):
    """BaseFDTD11.Material (:136-161)."""
    V.epsilon = np.ones(P.Nz)
    V.mu = np.ones(P.Nz)
    F, R = int(P.materialFrontEdge), int(P.materialRearEdge)
    for kj in range(P.Nz):
        if F <= kj < R:
            V.epsilon[kj], V.mu[kj] = P.epsRe, P.muRe
            V.UpExHcompsCo[kj], V.UpExSelf[kj] = P.eHcompsCo, P.eSelfCo
            V.UpHyEcompsCo[kj], V.UpHySelf[kj] = P.hEcompsCo, P.hSelfCo
        else:
            V.epsilon[kj] = V.mu[kj] = 1
            V.UpExHcompsCo[kj] = V.UpExSelf[kj] = V.UpHyEcompsCo[kj] = V.UpHySelf[kj] = 1
    return V.epsilon, V.mu, V.UpExHcompsCo, V.UpExSelf, V.UpHyEcompsCo, V.UpHySelf


def UpdateCoef(V, P):
    """BaseFDTD11.UpdateCoef (:168-182): slab coefficients divided by eps/mu."""
    UpHyMat = np.ones(P.Nz + 1) * ((1 / P.CharImp) * P.courantNo)
    UpExMat = np.ones(P.Nz + 1) * (P.CharImp * P.courantNo)
    for k in range(P.materialFrontEdge, P.materialRearEdge):
        UpExMat[k] = UpExMat[k] / V.epsilon[k]
        UpHyMat[k] = UpHyMat[k] / V.mu[k]
    return UpHyMat, UpExMat


def CPML_FieldInit(V, P, C_V, C_P):
    """BaseFDTD11.CPML_FieldInit (:195-219)."""
    n = P.Nz + 1
    C_V.kappa_Ex, C_V.kappa_Hy = np.ones(n), np.ones(n)
    for name in ("psi_Ex", "psi_Hy", "alpha_Ex", "alpha_Hy", "sigma_Ex", "sigma_Hy", "beX", "bmY",
                 "ceX", "cmY", "eLoss_CPML", "mLoss_CPML", "Ca", "Cb", "Cc", "C1", "C2", "C3"):
        setattr(C_V, name, np.zeros(n))
    C_V.den_Exdz = np.ones(len(V.Ex))
    C_V.den_Hydz = np.ones(len(V.Ex))
    return C_V


def CPML_ScalingCalc(V, P, C_V, C_P):
    """BaseFDTD11.CPML_ScalingCalc (:222-272): graded kappa/sigma/alpha; H aliases E."""
    W = P.pmlWidth
    n = len(V.Ex)
    nz = np.arange(0, W, 1)
    kap = 1 + (C_P.kappaMax - 1) * ((W - nz) / W) ** C_P.r_scale
    sig = C_P.sigmaOpt * ((W - nz) / W) ** C_P.r_scale
    alp = C_P.alphaMax * ((nz + 1) / W) ** C_P.r_a_scale
    C_V.kappa_Ex[:W], C_V.kappa_Ex[n - W:n] = kap, kap[::-1]
    C_V.sigma_Ex[:W], C_V.sigma_Ex[n - W:n] = sig, sig[::-1]
    C_V.alpha_Ex[:W], C_V.alpha_Ex[n - W:n] = alp, alp[::-1]
    C_V.kappa_Hy = C_V.kappa_Ex
    C_V.sigma_Hy = C_V.sigma_Ex
    C_V.alpha_Hy = C_V.alpha_Ex
    return C_V.sigma_Ex, C_V.sigma_Hy, C_V.alpha_Ex, C_V.alpha_Hy, C_V.kappa_Ex, C_V.kappa_Hy


def _rc(sigma, kappa, alpha, P, extra=1.0):
    b = np.exp(-((sigma * P.delT / (kappa * P.permit_0)) + ((alpha * P.delT) / P.permit_0)))
    c = (b - 1) * sigma / ((sigma * kappa + alpha * kappa * kappa) * extra) if extra != 1.0 else \
        (b - 1) * sigma / ((sigma * kappa + alpha * kappa * kappa))
    return b, c


def CPML_Ex_RC_Define(V, P, C_V, C_P):
    """BaseFDTD11.CPML_Ex_RC_Define (:274-286)."""
    W, n = P.pmlWidth, len(V.Ex)
    for sl in (slice(0, W), slice(n - W, n)):
        C_V.beX[sl], C_V.ceX[sl] = _rc(C_V.sigma_Ex[sl], C_V.kappa_Ex[sl], C_V.alpha_Ex[sl], P)
    return C_V.beX, C_V.ceX


def CPML_HY_RC_Define(V, P, C_V, C_P):
    """BaseFDTD11.CPML_HY_RC_Define (:288-296): as E with an extra /dz on cmY."""
    W, n = P.pmlWidth, len(V.Hy)
    for sl in (slice(0, W), slice(n - W, n)):
        C_V.bmY[sl], C_V.cmY[sl] = _rc(C_V.sigma_Hy[sl], C_V.kappa_Hy[sl], C_V.alpha_Hy[sl], P, P.dz)
    return C_V.bmY, C_V.cmY


def CPML_Ex_Update_Coef(V, P, C_V, C_P):
    """BaseFDTD11.CPML_Ex_Update_Coef (:299-318): Cb on [0,W) and (n-W, n-1]."""
    betaE = (0.5 * V.plasmaFreqE * V.plasmaFreqE * P.permit_0 * P.delT) / (1 + 0.5 * V.gammaE * P.delT)
    a = (2 * P.permit_0 - betaE * P.delT) / (2 * P.permit_0 + betaE * P.delT)
    n = len(V.Hy)
    for nz in list(range(0, P.pmlWidth)) + list(range(n - 1, n - P.pmlWidth, -1)):
        C_V.eLoss_CPML[nz] = (C_V.sigma_Ex[nz] * P.delT) / (2 * P.permit_0)
        C_V.Ca[nz] = a
        C_V.Cb[nz] = V.UpExHcompsCo[nz] * V.UpExMat[nz]
        C_V.Cc[nz] = P.delT / ((1 + C_V.eLoss_CPML[nz]) * P.permit_0)
    return C_V.eLoss_CPML, C_V.Ca, C_V.Cb, C_V.Cc


def CPML_Hy_Update_Coef(V, P, C_V, C_P):
    """BaseFDTD11.CPML_Hy_Update_Coef (:321-330): C2 = dt/mu0 on the same cells as Cb."""
    n = len(V.Hy)
    for nz in list(range(0, P.pmlWidth)) + list(range(n - 1, n - P.pmlWidth, -1)):
        C_V.C1[nz] = 1
        C_V.C2[nz] = P.delT / P.permea_0
        C_V.C3[nz] = P.delT / ((1 + C_V.mLoss_CPML[nz]) * P.permea_0)
    return C_V.mLoss_CPML, C_V.C1, C_V.C2, C_V.C3


def denominators(V, P, C_V, C_P):
    """BaseFDTD11.denominators (:332-357): 1/kappa, mirrored indices on the right."""
    n = len(V.Hy)
    jj = P.pmlWidth
    for j in range(n):
        if j <= P.pmlWidth and P.CPMLXm:
            C_V.den_Hydz[j] = 1 / C_V.kappa_Hy[j]
        elif j >= n - P.pmlWidth and P.CPMLXp:
            C_V.den_Hydz[j] = 1 / C_V.kappa_Hy[jj]
            jj -= 1
        else:
            C_V.den_Hydz[j] = 1.0
    jj = P.pmlWidth - 1
    for j in range(n):
        if j <= P.pmlWidth and P.CPMLXm:
            C_V.den_Exdz[j] = 1 / C_V.kappa_Ex[j]
        elif j >= n - P.pmlWidth and P.CPMLXp:
            C_V.den_Exdz[j] = 1 / C_V.kappa_Ex[jj]
            jj -= 1
        else:
            C_V.den_Exdz[j] = 1.0
    return C_V.den_Exdz, C_V.den_Hydz


def ade_coefficients(V, P):
    """A, B, C of ADE_PolarisationCurrent_Ex (BaseFDTD11.py:620-626)."""
    D = (1 / P.delT ** 2) + (V.gammaE / (2 * P.delT))
    A = ((2 / P.delT ** 2) - V.omega_0E ** 2) / D
    B = ((V.gammaE / (2 * P.delT)) - 1 / P.delT ** 2) / D
    C = (P.permit_0 * V.plasmaFreqE ** 2) / D
    return A, B, C


def cubic_coefficients(V, P):
    """cub, qua, one of AcubicFinder (BaseFDTD11.py:803-810)."""
    epsNum = V.plasmaFreqE * V.plasmaFreqE
    epsDom = (V.omega_0E * V.omega_0E - (2 * np.pi * P.freq_in * 2 * np.pi * P.freq_in)
              + 1j * V.gammaE * 2 * np.pi * P.freq_in)
    epsilon = 1 + epsNum / epsDom
    cub = (V.alpha3 * V.chi3Stat) ** 2
    qua = 2 * np.real(V.alpha3 * epsilon * V.chi3Stat)
    one = np.abs(epsilon) ** 2
    return float(cub), float(qua), float(one)


def AnalyticalReflectionE(V, P):
    """BaseFDTD11.AnalyticalReflectionE (:882-923): Fresnel |(n-1)/(n+1)|, n = Re sqrt(eps)."""
    epsNum = V.plasmaFreqE * V.plasmaFreqE
    epsDom = (V.omega_0E * V.omega_0E - (2 * np.pi * P.freq_in * 2 * np.pi * P.freq_in)
              + 1j * V.gammaE * 2 * np.pi * P.freq_in)
    epsilon = 1 + epsNum / epsDom
    refr2 = np.real(np.sqrt(epsilon))
    return abs((refr2 - 1) / (1 + refr2))


# ---------------------------------------------------------------------------
# per-step kernels: one fused device step each (compatibility path)
# ---------------------------------------------------------------------------
def _single(op):
    def fn(V, P, C_V=None, C_P=None, counts=0):
        from fdtd_pic_amd import single_step
        return single_step.apply(op, V, P, C_V, counts)
    fn.__name__ = op
    fn.__doc__ = f"BaseFDTD11.{op} executed on the GPU (fdtd_pic_amd.single_step)."
    return fn


ADE_ExUpdate = _single("ADE_ExUpdate")
ADE_HyUpdate = _single("ADE_HyUpdate")
ADE_DxUpdate = _single("ADE_DxUpdate")
ADE_ExCreate = _single("ADE_ExCreate")
CPML_Psi_e_Update = _single("CPML_Psi_e_Update")
CPML_Psi_m_Update = _single("CPML_Psi_m_Update")


def ADE_PolarisationCurrent_Ex(V, P, C_V, C_P, counts):
    from fdtd_pic_amd import single_step
    return single_step.apply("ADE_PolarisationCurrent_Ex", V, P, C_V, counts)


def ADE_TempPolCurr(V, P, C_V, C_P):
    """BaseFDTD11.ADE_TempPolCurr (:487-538): history shift of P and E by copy
    (the reference goes through lists for the same effect) and of Dx, Hy, Jx,
    psi by reference; returns the reference's 12-tuple.  The reference's
    breakpoint() trap on an unchanged polarisation (:535-537) is not kept."""
    V.tempTempVarPol = np.array(V.tempVarPol, copy=True)
    V.tempVarPol = np.array(V.polarisationCurr, copy=True)
    V.tempTempVarDx = V.tempVarDx
    V.tempVarDx = V.Dx
    V.tempTempVarE = np.array(V.tempVarE, copy=True)
    V.tempVarE = np.array(V.Ex, copy=True)
    V.tempTempVarHy = V.tempVarHy
    V.tempVarHy = V.Hy
    V.tempTempVarJx = V.tempVarJx
    V.tempVarJx = V.Jx
    C_V.tempTempVarPsiEx = C_V.tempVarPsiEx
    C_V.tempVarPsiEx = C_V.psi_Ex
    C_V.tempTempVarPsiHy = C_V.tempVarPsiHy
    C_V.tempVarPsiHy = C_V.psi_Hy
    return (V.tempTempVarPol, V.tempVarPol, V.tempVarE, V.tempTempVarE, V.tempTempVarHy, V.tempVarHy,
            V.tempTempVarJx, V.tempVarJx, C_V.tempTempVarPsiEx, C_V.tempVarPsiEx, C_V.tempTempVarPsiHy,
            C_V.tempVarPsiHy)


def AcubicFinder(V, P):
    from fdtd_pic_amd import single_step
    return single_step.apply("AcubicFinder", V, P, None, 0)


def NonLinExUpdate(V, P):
    from fdtd_pic_amd import single_step
    return single_step.apply("NonLinExUpdate", V, P, None, 0)
