"""Oracle: NumPy FP64 restatement of the reference integrators.  TEST INFRASTRUCTURE ONLY.

Every function restates the reference function it names, statement by
statement, vectorised over cells where the reference loop carries no
dependence between iterations (each such loop reads only the other field).
Reference quirks are kept on purpose, because the product must reproduce
them to 1e-10:

* float32 jitclass members: Params.courantNo/Amplitude/Periods/epsRe/muRe
  (MasterController.py:234,247,273-279), Variables.alpha3 (:129),
  CPML_Params.kappaMax/r_scale/r_a_scale/alphaMax (:341-347).  Duplicate
  spec keys resolve last-wins (timeSteps int32, lamMin float64, ...).
* CharImp hard-coded 376.730313668 (:293) while eps0/mu0 come from scipy.
* cumulative spatialStab rescale of plasmaFreqE per pass (Solver_Engine.py:286).
* CPML: E side subtracts psi (BaseFDTD11.py:371,375); Cb/C2 hole at
  N+1-W (:312,326); H profiles alias E profiles (:266-268); cmY carries /dz.
* NL: cubic in x=|E|^2 with the complex Lorentz eps and alpha3
  (BaseFDTD11.py:803-820) but E = Dx/(eps0*sqrt(1.2) + eps0*chi3*A) (:858-877).

The reference's missing `CubicEquationSolver.CubicSolver` class (:836) is
replaced by the vendored `solve(a,b,c,d)` on the real parts of the stored
complex coefficients; `Tests.*` modules are taken to be the root-level copies.
"""
import math
import types

import numpy as np

from .cubic import root0_1728

# scipy 1.15.3 (CODATA 2022) values as used by MasterController.py:291-292
EPS0 = 8.8541878188e-12
MU0 = 1.25663706127e-06
CHAR_IMP = 376.730313668          # MasterController.py:293
C0 = 299792458.0                  # MasterController.py:294


def f32(x):
    return float(np.float32(x))


class RefExit(SystemExit):
    """Where the reference prints and calls sys.exit()."""


# ---------------------------------------------------------------------------
# L0: parameter construction
# ---------------------------------------------------------------------------
def env_setup(newFreq_in, domainSize, minim=400, maxim=600, VExists=False, V=None, P=None,
              nonLinMed=False, LorMed=False):
    """Environment_Setup.envSetup (Environment_Setup.py:19-166) -> 20-tuple."""
    pi = np.pi
    epsilon = None
    if VExists:                                                       # :23-27
        epsNum = V.plasmaFreqE * V.plasmaFreqE
        epsDom = (V.omega_0E * V.omega_0E - (2 * pi * P.freq_in * 2 * pi * P.freq_in)
                  + 1j * V.gammaE * 2 * pi * P.freq_in)
        epsilon = 1 + epsNum / epsDom
    freq_in = newFreq_in
    c0 = 299792458.0
    lamMin = c0 / freq_in
    if VExists:                                                       # :38-46
        Nlam = int(60 * (np.real(epsilon)) ** 1.05)
        if nonLinMed:
            Nlam = int(200 * (np.real(epsilon)) ** 1.05)
    else:
        Nlam = 400
        if nonLinMed:
            Nlam = 350
    dz = lamMin / Nlam
    delT = (dz / c0) * 0.95                                           # :49
    period = 1 / freq_in
    courantNo = (c0 * delT) / dz
    if courantNo > 3 or courantNo < 0:
        raise RefExit("courantNo is unstable")
    pmlWidth = 6 * int(lamMin / dz)                                   # :62
    if pmlWidth >= 12000:
        raise RefExit("pmlWidth too big")
    dimen = 1
    nonInt = False
    Nz = int(domainSize / dz) + 2 * dimen * pmlWidth                  # :73
    if minim == maxim:
        minim -= 1
    N = None
    timeSteps = None
    for N in range(minim, maxim):                                     # :84-95
        check = (freq_in * N) / (1 / delT)
        if int(check) - check == 0:
            timeSteps = N
            break
        elif N == maxim - 1:
            nonInt = True
    if nonInt:                                                        # :101-110
        checkNear = (freq_in * minim) / (1 / delT)
        N = 0
        for N in range(minim, maxim):
            dummyCheck = (freq_in * N) / (1 / delT)
            if int(dummyCheck) - dummyCheck < int(checkNear) - checkNear:
                checkNear = dummyCheck
    if N >= minim:                                                    # :113-116
        timeSteps = N
    else:
        timeSteps = minim
    timeSteps += int(timeSteps * (Nlam / 200))                        # :118
    if timeSteps >= 2 ** 15:
        raise RefExit("timeSteps too large")
    nzsrcFromPml = int(0.05 / dz)
    if nzsrcFromPml >= Nz * 0.65:
        raise RefExit("src is too far into domain")
    nzsrc = nzsrcFromPml + pmlWidth
    if nzsrc - 10 <= pmlWidth:
        raise RefExit("The probe for fft is in the PML region")
    MaterialDistFromPml = int(0.1 / dz)
    materialFrontEdge = MaterialDistFromPml + pmlWidth
    materialRearEdge = Nz - 1
    if materialRearEdge - materialFrontEdge < 10:
        raise RefExit("width is too small or negative")
    if MaterialDistFromPml >= domainSize / dz:
        raise RefExit("Material starts in CPML region")
    if materialFrontEdge <= nzsrc:
        raise RefExit("Source is inside material")
    x1Loc = materialFrontEdge - 20
    x2Loc = nzsrc - 100
    eLoss = 0
    mLoss = 0
    eSelfCo = (1 - eLoss) / (1 + eLoss)
    eHcompsCo = 1 / (1 + eLoss)
    hSelfCo = (1 - mLoss) / (1 + mLoss)
    hEcompsCo = 1 / (1 + mLoss)
    return (Nz, timeSteps, eLoss, mLoss, eSelfCo, eHcompsCo, hSelfCo, hEcompsCo, x1Loc, x2Loc,
            materialFrontEdge, materialRearEdge, pmlWidth, nzsrc, lamMin, dz, delT, courantNo,
            period, Nlam)


def mat_setup(V, P, C_P, C_V, newFreq_in, domainSize, minim=400, maxim=800):
    """Material_Def.matSetup (Material_Def.py:21-120) -> 19-tuple (no Nlam)."""
    freq_in = newFreq_in
    c0 = 299792458.0
    lamMin = (c0 / freq_in) * 0.1
    Nlam = 20
    dz = lamMin / Nlam
    delT = 0.99 / (c0 * (1 / dz))
    period = 1 / freq_in
    courantNo = (c0 * delT) / dz
    if courantNo > 1 or courantNo < 0:
        raise RefExit("courantNo is unstable")
    pmlWidth = 150
    domainSize = 600
    Nz = domainSize + 2 * 1 * pmlWidth
    nonInt = False
    N = None
    for N in range(minim, maxim):
        check = (freq_in * N) / (1 / delT)
        if int(check) - check == 0:
            break
        elif N == maxim - 1:
            nonInt = True
    if nonInt:
        checkNear = (freq_in * minim) / (1 / delT)
        for N in range(minim, maxim):
            dummyCheck = (freq_in * N) / (1 / delT)
            if int(dummyCheck) - dummyCheck < int(checkNear) - checkNear:
                checkNear = dummyCheck
    timeSteps = N                                                     # :79
    if timeSteps >= 10000:
        raise RefExit("timeSteps too large")
    nzsrcFromPml = 100
    nzsrc = nzsrcFromPml + pmlWidth
    materialFrontEdge = 300 + pmlWidth
    materialRearEdge = Nz - pmlWidth - 1
    x1Loc = nzsrc + int((materialFrontEdge - nzsrc) / 2)
    x2Loc = nzsrc - (nzsrcFromPml + 10) + 50
    eLoss = 0
    mLoss = 0
    return (Nz, timeSteps, eLoss, mLoss, (1 - eLoss) / (1 + eLoss), 1 / (1 + eLoss),
            (1 - mLoss) / (1 + mLoss), 1 / (1 + mLoss), x1Loc, x2Loc, materialFrontEdge,
            materialRearEdge, pmlWidth, nzsrc, lamMin, dz, delT, courantNo, period)


def make_params(setup, MORmode, domainSize, freq_in, delayMOR, LorentzMed=False, nonLinMed=False,
                SineCont=False, Gaussian=False, TFSF=False):
    """MasterController.Params (MasterController.py:217-336) with spec types applied."""
    (Nz, timeSteps, eLoss, mLoss, eSelfCo, eHcompsCo, hSelfCo, hEcompsCo, x1Loc, x2Loc, F, R, W,
     nzsrc, lamMin, dz, delT, courantNo, period, Nlam) = setup
    P = types.SimpleNamespace()
    P.permit_0, P.permea_0, P.CharImp, P.c0 = EPS0, MU0, CHAR_IMP, C0
    P.freq_in = float(freq_in)
    P.lamMin = float(lamMin)
    P.Nlam = float(Nlam)
    P.dz = float(dz)
    P.delT = float(delT)
    P.courantNo = f32(courantNo)
    P.materialFrontEdge, P.materialRearEdge = int(F), int(R)
    P.Nz, P.timeSteps = int(Nz), int(timeSteps)
    P.x1Loc, P.x2Loc, P.nzsrc = int(x1Loc), int(x2Loc), int(nzsrc)
    P.period = float(period)
    P.eLoss, P.eSelfCo, P.eHcompsCo = float(eLoss), float(eSelfCo), float(eHcompsCo)
    P.mLoss, P.hSelfCo, P.hEcompsCo = float(mLoss), float(hSelfCo), float(hEcompsCo)
    P.pmlWidth = int(W)
    P.domainSize = int(domainSize)
    P.MORmode, P.delayMOR = bool(MORmode), int(delayMOR)
    P.CPMLXp = P.CPMLXm = True
    P.TFSF, P.SineCont, P.Gaussian, P.Ricker = bool(TFSF), bool(SineCont), bool(Gaussian), False
    P.Amplitude = f32(1.0)
    P.Periods = f32(1.0)
    P.LorentzMed, P.nonLinMed, P.FreeSpace = bool(LorentzMed), bool(nonLinMed), True
    P.epsRe = f32(1.0)
    P.muRe = f32(1.0)
    P.vidMake, P.vidInterval, P.atten, P.julia, P.testMode = True, 50, False, False, False
    return P


def make_variables(Nz, timeSteps, interval, attenAmnt):
    """MasterController.Variables (MasterController.py:146-210): the members the loop reads."""
    V = types.SimpleNamespace()
    n = Nz + 1
    for name in ("UpHyMat", "UpExMat", "UpExHcompsCo", "UpExSelf", "UpHySelf", "UpHyEcompsCo",
                 "epsilon", "mu"):
        setattr(V, name, np.ones(n))
    for name in ("Ex", "Hy", "Dx", "Jx", "polarisationCurr", "tempVarPol", "tempTempVarPol",
                 "tempVarE", "tempTempVarE", "Acubic"):
        setattr(V, name, np.zeros(n))
    rows = int(timeSteps / interval)
    V.Ex_History = np.zeros((rows, n))
    for name in ("x1ColBe", "x1ColAf", "Port1", "Port2"):
        setattr(V, name, np.zeros(timeSteps))
    V.plasmaFreqE = np.sqrt(1.5 * (2 * np.pi * 20e9) ** 2)           # :177
    V.gammaE = 2 * np.pi * 20e9 * 0.1                                 # :178
    V.omega_0E = 2 * np.pi * 20e9                                     # :179
    V.attenAmnt = int(attenAmnt)
    V.x1Atten = np.zeros((attenAmnt, timeSteps))
    V.alpha3 = f32(0.7)                                               # :196 (float32)
    V.chi1Stat = np.sqrt(1.2) - 1
    V.chi3Stat = 1e-3
    V.cubPoly = np.zeros((n, 4), dtype=np.complex128)
    return V


def make_cpml_params(dz):
    """MasterController.CPML_Params (MasterController.py:351-360)."""
    C_P = types.SimpleNamespace()
    C_P.kappaMax = f32(1)
    C_P.r_scale = f32(4)
    C_P.r_a_scale = f32(1)
    C_P.sigmaEMax = 0.5 * (0.8 * 1 / (dz * (MU0 / EPS0) ** 0.5))
    C_P.sigmaHMax = C_P.sigmaEMax
    C_P.sigmaOpt = C_P.sigmaEMax
    C_P.alphaMax = f32(0.05)
    return C_P


def make_cpml_variables(Nz, timeSteps):
    C_V = types.SimpleNamespace()
    for name in ("psi_Ex", "psi_Hy"):
        setattr(C_V, name, np.zeros(Nz + 1))
    return C_V


# ---------------------------------------------------------------------------
# genericStability.spatialStab (genericStability.py:12-62)
# ---------------------------------------------------------------------------
def spatial_stab(fullTime, fullSpace, spaceStep, frequency, timeStep, plasmaFreq, resonantFreq, gamma):
    fS = fullSpace
    sp = spaceStep
    los = np.pi / 2
    frq = frequency
    rad = 2 * np.pi * frq
    twoPi = rad / frq
    c0 = 299792458.0
    tim = timeStep
    if rad * tim > np.pi / 2 and frq <= 5e9:
        raise RefExit("unstable timestep")
    sw = np.sinc(np.pi * frq * tim)
    pf = plasmaFreq
    rf = resonantFreq
    es = (pf ** 2) / (rf ** 2) - 1
    gam = gamma
    sqN = rad ** 2 * sw ** 2 - es * rf ** 2 * np.cos(rad * tim) + 1.0j * gam * rad * sw
    sqD = rad ** 2 * sw ** 2 - rf ** 2 * np.cos(rad * tim) + 1.0j * gam * rad * sw
    arg = (rad / c0) * (sp / 2) * sw * np.sqrt(sqN / sqD)
    ans = (2 / sp) * np.arcsin(arg)
    kNum = abs(ans)
    epsilon = 1 + (pf ** 2) / (rf ** 2 - rad ** 2 - 1j * gam * rad)
    refr = np.sqrt(abs(np.real(epsilon)))
    fix = (c0 * tim * np.sin((kNum * refr * sp) / 2)) / (refr * sp * np.sin((kNum * c0 * tim) / 2))
    pf = np.sqrt(abs(fix)) * pf
    lamCont = c0 / frq
    lamDisc = twoPi / kNum
    diff = abs(lamCont - lamDisc)
    if kNum * sp > los:
        raise RefExit("unstable, wave is not resolved")
    if kNum * sp * fS < (5 * lamDisc):
        raise RefExit("unstable because domain too small")
    return lamCont, lamDisc, diff, pf, fix


# ---------------------------------------------------------------------------
# L1 set-up kernels
# ---------------------------------------------------------------------------
def field_init(V, P):
    """BaseFDTD11.FieldInit (BaseFDTD11.py:36-82)."""
    if P.timeSteps > 2 ** 17:
        raise RefExit("timeSteps max too large")
    if P.Nz > 25000:
        raise RefExit("Grid size too big")
    if P.Nz == 0 or P.timeSteps == 0 or P.Nz < 0 or P.timeSteps < 0:
        raise RefExit("bad grid/timeSteps")
    n = P.Nz + 1
    V.Ex, V.Hy, V.polarisationCurr, V.Dx = np.zeros(n), np.zeros(n), np.zeros(n), np.zeros(n)
    V.tempVarPol, V.tempTempVarPol = np.zeros(n), np.zeros(n)
    V.tempVarE, V.tempTempVarE = np.zeros(n), np.zeros(n)


def empty_space_calc(V, P):
    """BaseFDTD11.EmptySpaceCalc (:124-132)."""
    V.UpHyMat = np.ones(P.Nz + 1)
    V.UpExMat = np.ones(P.Nz + 1)
    V.UpExMat *= P.CharImp * P.courantNo
    V.UpHyMat *= (1 / P.CharImp) * P.courantNo


def material(V, P):
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


def update_coef(V, P):
    """BaseFDTD11.UpdateCoef (:168-182)."""
    UpHyMat = np.ones(P.Nz + 1) * ((1 / P.CharImp) * P.courantNo)
    UpExMat = np.ones(P.Nz + 1) * (P.CharImp * P.courantNo)
    for k in range(P.materialFrontEdge, P.materialRearEdge):
        UpExMat[k] = UpExMat[k] / V.epsilon[k]
        UpHyMat[k] = UpHyMat[k] / V.mu[k]
    V.UpHyMat, V.UpExMat = UpHyMat, UpExMat


def bound_cond_manager(V, P, C_V, C_P):
    """CPML_FieldInit + boundCondManager (BaseFDTD11.py:195-357, Solver_Engine.py:70-87)."""
    n = P.Nz + 1
    W = P.pmlWidth
    C_V.psi_Ex, C_V.psi_Hy = np.zeros(n), np.zeros(n)
    kappa, sigma, alpha = np.ones(n), np.zeros(n), np.zeros(n)
    beX, ceX, bmY, cmY = np.zeros(n), np.zeros(n), np.zeros(n), np.zeros(n)
    Cb, C2 = np.zeros(n), np.zeros(n)
    den_E, den_H = np.ones(n), np.ones(n)
    # CPML_ScalingCalc (:223-272)
    nz = np.arange(0, W, 1)
    kap_n = 1 + (C_P.kappaMax - 1) * ((W - nz) / W) ** C_P.r_scale
    sig_n = C_P.sigmaOpt * ((W - nz) / W) ** C_P.r_scale
    alp_n = C_P.alphaMax * ((nz + 1) / W) ** C_P.r_a_scale
    kappa[:W], sigma[:W], alpha[:W] = kap_n, sig_n, alp_n
    kappa[n - W:n], sigma[n - W:n], alpha[n - W:n] = kap_n[::-1], sig_n[::-1], alp_n[::-1]
    # H profiles alias E profiles (:266-268)
    # CPML_Ex_RC_Define (:275-286), CPML_HY_RC_Define (:289-296)
    for sl in (slice(0, W), slice(n - W, n)):
        beX[sl] = np.exp(-((sigma[sl] * P.delT / (kappa[sl] * P.permit_0))
                           + ((alpha[sl] * P.delT) / P.permit_0)))
        ceX[sl] = (beX[sl] - 1) * sigma[sl] / (sigma[sl] * kappa[sl] + alpha[sl] * kappa[sl] * kappa[sl])
        bmY[sl] = beX[sl]
        cmY[sl] = (bmY[sl] - 1) * sigma[sl] / ((sigma[sl] * kappa[sl]
                                                + alpha[sl] * kappa[sl] * kappa[sl]) * P.dz)
    # CPML_Ex_Update_Coef / CPML_Hy_Update_Coef (:299-330): hole at n-W
    idx = list(range(0, W)) + list(range(n - 1, n - W, -1))
    for k in idx:
        Cb[k] = V.UpExHcompsCo[k] * V.UpExMat[k]
        C2[k] = P.delT / P.permea_0
    # denominators (:333-357)
    jj = W
    for j in range(n):
        if j <= W and P.CPMLXm:
            den_H[j] = 1 / kappa[j]
        elif j >= n - W and P.CPMLXp:
            den_H[j] = 1 / kappa[jj]
            jj -= 1
    jj = W - 1
    for j in range(n):
        if j <= W and P.CPMLXm:
            den_E[j] = 1 / kappa[j]
        elif j >= n - W and P.CPMLXp:
            den_E[j] = 1 / kappa[jj]
            jj -= 1
    C_V.kappa_Ex = C_V.kappa_Hy = kappa
    C_V.sigma_Ex = C_V.sigma_Hy = sigma
    C_V.alpha_Ex = C_V.alpha_Hy = alpha
    C_V.beX, C_V.ceX, C_V.bmY, C_V.cmY, C_V.Cb, C_V.C2 = beX, ceX, bmY, cmY, Cb, C2
    C_V.den_Exdz, C_V.den_Hydz = den_E, den_H


def smooth_turn_on(P, tempfreq=0):
    """BaseFDTD11.SmoothTurnOn (:104-120)."""
    frq = tempfreq if tempfreq != 0 else P.freq_in
    ppw = P.c0 / (frq * P.dz)
    Exs, Hys = [], []
    for timer in range(P.timeSteps):
        if timer * P.delT < P.period * P.Periods:
            Exs.append(float(np.sin(2.0 * np.pi / ppw * (P.courantNo * timer))))
            Hys.append(float(np.sin(2.0 * np.pi / ppw * (P.courantNo * (timer + 1)))))
        else:
            Exs.append(0)
            Hys.append(0)
    return Exs, Hys


def gaussian_src(P):
    """BaseFDTD11.Gaussian (:86-96)."""
    t = np.arange(P.timeSteps)
    fc = 200
    tau = fc * 2.2
    gaussArg = ((t - tau) * (t - tau)) / (fc * fc) * np.cos(2 * np.pi * P.freq_in * (t - tau))
    out = np.exp(-gaussArg) * 2
    return P.Amplitude * out / P.courantNo


def source_manager(P):
    """Solver_Engine.SourceManager (Solver_Engine.py:89-124)."""
    if P.SineCont:
        Exs, Hys1 = smooth_turn_on(P)
        Exp, Hyp1 = np.zeros(P.timeSteps), np.zeros(P.timeSteps)
        if P.nonLinMed:
            Exp, Hyp1 = smooth_turn_on(P, tempfreq=P.freq_in * 0.8)
            Exp = np.asarray(Exp) * P.courantNo
            Hyp1 = np.asarray(Hyp1) * P.courantNo
            Exp *= 0.1
            Hyp1 *= 0.01
        Exs = np.asarray(Exs) * P.courantNo + Exp
        Hys1 = np.asarray(Hys1) * P.courantNo + Hyp1
        if P.TFSF:
            Hys1 = Hys1 * (1 / P.CharImp)
        return Exs, Hys1
    if P.Gaussian:
        Exs = gaussian_src(P)
        Hys = gaussian_src(P) if P.TFSF else np.zeros(len(Exs))
        return Exs, Hys
    return [], []


def sig_mod(P, sig, AmpMod=1, tau=14.6e-10):
    """Solver_Engine.Sig_Mod / SechArr (Solver_Engine.py:126-140)."""
    trang = np.arange(len(sig)) * P.delT
    omega_mod = 2 * np.pi * (1 / tau)
    a = omega_mod * trang
    a = np.where(a > 50, 50, a)
    return (AmpMod + sig) * (1 / np.cosh(a))


# ---------------------------------------------------------------------------
# L1 per-step kernels
# ---------------------------------------------------------------------------
def ade_ex_update(V, P, C_V):
    # BaseFDTD11.ADE_ExUpdate (:663-669)
    V.Ex[1:] = V.Ex[1:] + (V.Hy[1:] - V.Hy[:-1] - V.Jx[1:]) * V.UpExMat[1:] * C_V.den_Exdz[1:]


def cpml_psi_e(V, P, C_V):
    # BaseFDTD11.CPML_Psi_e_Update (:364-376)
    n = P.Nz + 1
    W = P.pmlWidth
    for lo, hi in ((1, W), (n - W, n)):
        s = slice(lo, hi)
        sm = slice(lo - 1, hi - 1)
        C_V.psi_Ex[s] = C_V.beX[s] * C_V.psi_Ex[s] + C_V.ceX[s] * (V.Hy[s] - V.Hy[sm])
        V.Ex[s] = V.Ex[s] - C_V.Cb[s] * C_V.psi_Ex[s]


def cpml_psi_m(V, P, C_V):
    # BaseFDTD11.CPML_Psi_m_Update (:381-393)
    n = P.Nz + 1
    W = P.pmlWidth
    for lo, hi in ((1, W), (n - W, n - 1)):
        s = slice(lo, hi)
        sp = slice(lo + 1, hi + 1)
        C_V.psi_Hy[s] = C_V.bmY[s] * C_V.psi_Hy[s] + C_V.cmY[s] * (V.Ex[sp] - V.Ex[s])
        V.Hy[s] = V.Hy[s] + C_V.C2[s] * C_V.psi_Hy[s]


def ade_hy_update(V, P, C_V):
    # BaseFDTD11.ADE_HyUpdate (:640-656)
    N = P.Nz
    V.Hy[1:N] = V.Hy[1:N] * V.UpHySelf[1:N] + (V.Ex[2:N + 1] - V.Ex[1:N]) * V.UpHyMat[1:N] * C_V.den_Hydz[1:N]


def ade_dx_update(V, P, C_V):
    # BaseFDTD11.ADE_DxUpdate (:750-760)
    F, R = P.materialFrontEdge, P.materialRearEdge
    V.Dx[F:R] = V.Dx[F:R] + (V.Hy[F:R] - V.Hy[F - 1:R - 1]) * (P.delT / P.dz) * C_V.den_Exdz[F:R]


def ade_ex_create(V, P):
    # BaseFDTD11.ADE_ExCreate (:712-725)
    F, R = P.materialFrontEdge, P.materialRearEdge
    V.Ex[F:R] = (V.Dx[F:R] - V.polarisationCurr[F:R]) / P.permit_0


def ade_coefficients(V, P):
    """A, B, C of ADE_PolarisationCurrent_Ex (BaseFDTD11.py:620-626)."""
    D = (1 / P.delT ** 2) + (V.gammaE / (2 * P.delT))
    A = ((2 / P.delT ** 2) - V.omega_0E ** 2) / D
    B = ((V.gammaE / (2 * P.delT)) - 1 / P.delT ** 2) / D
    C = (P.permit_0 * V.plasmaFreqE ** 2) / D
    return A, B, C


def ade_temp_pol_curr(V):
    # BaseFDTD11.ADE_TempPolCurr (:487-538): history shift, copies
    V.tempTempVarPol = V.tempVarPol.copy()
    V.tempVarPol = V.polarisationCurr.copy()
    V.tempTempVarE = V.tempVarE.copy()
    V.tempVarE = V.Ex.copy()


def ade_polarisation(V, P):
    # BaseFDTD11.ADE_PolarisationCurrent_Ex (:609-633)
    A, B, C = ade_coefficients(V, P)
    F, R = P.materialFrontEdge, P.materialRearEdge
    V.polarisationCurr[F:R] = A * V.polarisationCurr[F:R] + B * V.tempTempVarPol[F:R] + C * V.Ex[F:R]


def cubic_coefficients(V, P):
    """cub, qua, one of AcubicFinder (BaseFDTD11.py:804-810)."""
    epsNum = V.plasmaFreqE * V.plasmaFreqE
    epsDom = (V.omega_0E * V.omega_0E - (2 * np.pi * P.freq_in * 2 * np.pi * P.freq_in)
              + 1j * V.gammaE * 2 * np.pi * P.freq_in)
    epsilon = 1 + epsNum / epsDom
    cub = (V.alpha3 * V.chi3Stat) ** 2
    qua = 2 * np.real(V.alpha3 * epsilon * V.chi3Stat)
    one = np.abs(epsilon) ** 2
    return float(cub), float(qua), float(one)


def acubic_finder(V, P):
    # AcubicFinder + Nonlin_Eqn_Setup + Nonlin_Cubic_Solver (:793-853)
    cub, qua, one = cubic_coefficients(V, P)
    F, R = P.materialFrontEdge, P.materialRearEdge
    for nz in range(F, R):
        d = -np.abs(V.Dx[nz] / P.permit_0) ** 2
        V.Acubic[nz] = root0_1728(cub, qua, one, d)


def nonlin_ex_update(V, P):
    # BaseFDTD11.NonLinExUpdate (:858-877)
    F, R = P.materialFrontEdge, P.materialRearEdge
    eps0 = P.permit_0
    epsilon = np.sqrt(1.2)
    V.Ex[F:R] = V.Dx[F:R] / (eps0 * np.real(epsilon) + eps0 * V.chi3Stat * V.Acubic[F:R])


def vid_make(V, P, counts):
    # Solver_Engine.vidMake (:57-68)
    shrunk = int(counts / P.vidInterval)
    if shrunk < len(V.Ex_History):
        V.Ex_History[shrunk] = V.Ex


def atten_probe(V, P, counts, granAttn=25):
    # Solver_Engine.probeSim(attenRead=True) (:31-38)
    inds = np.arange(P.materialFrontEdge, granAttn * V.attenAmnt + P.materialFrontEdge, granAttn)
    for i in range(len(V.x1Atten)):
        V.x1Atten[i][counts] = V.Ex[inds[i]]


# ---------------------------------------------------------------------------
# L2 integrators
# ---------------------------------------------------------------------------
def integrator_linlor(V, P, C_V, C_P, probeReadFinishBe, probeReadStartAf):
    """Solver_Engine.IntegratorLinLor1D (Solver_Engine.py:275-371)."""
    Exs = Hys = None
    for i in range(0, 2):
        field_init(V, P)
        empty_space_calc(V, P)
        bound_cond_manager(V, P, C_V, C_P)
        _, _, _, V.plasmaFreqE, _ = spatial_stab(P.timeSteps, P.Nz, P.dz, P.freq_in, P.delT,
                                                  V.plasmaFreqE, V.omega_0E, V.gammaE)
        Exs, Hys = source_manager(P)
        for counts in range(P.timeSteps):
            if i == 1:
                ade_temp_pol_curr(V)
                ade_polarisation(V, P)
            ade_ex_update(V, P, C_V)
            cpml_psi_e(V, P, C_V)
            V.Ex[P.nzsrc] += Exs[counts] / P.courantNo
            if P.TFSF:
                V.Hy[P.nzsrc - 1] -= Hys[counts] / P.courantNo
            ade_dx_update(V, P, C_V)
            ade_ex_create(V, P)
            ade_hy_update(V, P, C_V)
            cpml_psi_m(V, P, C_V)
            if counts > 0 and counts % P.vidInterval == 0 and i == 1:
                vid_make(V, P, counts)
            if i == 0:
                if counts <= probeReadFinishBe:
                    V.x1ColBe[counts] = V.Ex[P.x1Loc]
            else:
                if counts >= probeReadStartAf:
                    V.x1ColAf[counts] = V.Ex[P.x2Loc]
                    if P.atten:
                        atten_probe(V, P, counts)
    return V.Ex, V.Hy, Exs, Hys, C_V.psi_Ex, C_V.psi_Hy, V.x1ColBe, V.x1ColAf


def integrator_nl(V, P, C_V, C_P, probeReadFinishBe, probeReadStartAf):
    """Solver_Engine.IntegratorNL1D (Solver_Engine.py:220-271)."""
    field_init(V, P)
    empty_space_calc(V, P)
    bound_cond_manager(V, P, C_V, C_P)
    _, _, _, V.plasmaFreqE, _ = spatial_stab(P.timeSteps, P.Nz, P.dz, P.freq_in, P.delT,
                                              V.plasmaFreqE, V.omega_0E, V.gammaE)
    Exs, Hys = source_manager(P)
    for counts in range(P.timeSteps):
        ade_ex_update(V, P, C_V)
        cpml_psi_e(V, P, C_V)
        V.Ex[P.nzsrc] += Exs[counts] / P.courantNo
        if P.TFSF:
            V.Hy[P.nzsrc - 1] -= Hys[counts] / P.courantNo
        ade_dx_update(V, P, C_V)
        acubic_finder(V, P)
        nonlin_ex_update(V, P)
        ade_hy_update(V, P, C_V)
        cpml_psi_m(V, P, C_V)
        if counts > 0 and counts % P.vidInterval == 0:
            vid_make(V, P, counts)
        V.Port1[counts] = V.Ex[P.materialFrontEdge]
        V.Port2[counts] = V.Ex[P.materialRearEdge]
    return V.Ex, V.Hy, Exs, Hys, C_V.psi_Ex, C_V.psi_Hy, V.x1ColBe, V.x1ColAf


def integrator_free_space(V, P, C_V, C_P, probeReadFinishBe, probeReadStartAf):
    """Solver_Engine.IntegratorFreeSpace1D (Solver_Engine.py:142-215), julia=False."""
    Exs = Hys = None
    for i in range(0, 2):
        field_init(V, P)
        empty_space_calc(V, P)
        material(V, P)
        update_coef(V, P)
        bound_cond_manager(V, P, C_V, C_P)
        Exs, Hys = source_manager(P)
        tauIn = 1 / (P.freq_in / 5)
        Exs = sig_mod(P, Exs, tau=tauIn)
        Hys = sig_mod(P, Hys, AmpMod=1 / P.CharImp, tau=tauIn)
        for counts in range(P.timeSteps):
            ade_ex_update(V, P, C_V)
            cpml_psi_e(V, P, C_V)
            V.Ex[P.nzsrc] += Exs[counts] / P.courantNo
            if P.TFSF:
                V.Hy[P.nzsrc - 1] -= Hys[counts] / P.courantNo
            ade_hy_update(V, P, C_V)
            cpml_psi_m(V, P, C_V)
            if counts > 0 and counts % P.vidInterval == 0 and i == 1:
                vid_make(V, P, counts)
            if i == 0:
                if counts <= probeReadFinishBe:
                    V.x1ColBe[counts] = V.Ex[P.x1Loc]
            else:
                if counts >= probeReadStartAf:
                    V.x1ColAf[counts] = V.Ex[P.x2Loc]
                    if P.atten:
                        atten_probe(V, P, counts)
    return V.Ex, V.Hy, Exs, Hys, C_V.psi_Ex, C_V.psi_Hy, V.x1ColBe, V.x1ColAf


def controller(V, P, C_V, C_P):
    """MasterController.Controller (MasterController.py:451-469)."""
    probeReadFinishBe = int(P.timeSteps * 0.7)
    probeReadStartAf = int(P.timeSteps * 0.05)
    V.x1ColBe = np.zeros(P.timeSteps)
    V.x1ColAf = np.zeros(P.timeSteps)
    Exs = Hys = None
    if P.LorentzMed:
        out = integrator_linlor(V, P, C_V, C_P, probeReadFinishBe, probeReadStartAf)
    elif P.FreeSpace:
        out = integrator_free_space(V, P, C_V, C_P, probeReadFinishBe, probeReadStartAf)
    elif P.nonLinMed:
        out = integrator_nl(V, P, C_V, C_P, probeReadFinishBe, probeReadStartAf)
    else:
        return V, P, C_V, C_P, Exs, Hys
    V.Ex, V.Hy, Exs, Hys, C_V.psi_Ex, C_V.psi_Hy, V.x1ColBe, V.x1ColAf = out
    return V, P, C_V, C_P, Exs, Hys


def ref_tester(P, y):
    """TransformHandler.RefTester peak-bin amplitude (TransformHandler.py:34-65)."""
    Y = np.fft.fft(y)
    Ypow = (2 * np.abs(Y)) / P.timeSteps
    indMax = int(np.argmax(Ypow))
    if indMax == 0:
        raise RefExit("Could not find non-DC freq")
    return float(Ypow[indMax])


def analytical_reflection(V, P):
    """BaseFDTD11.AnalyticalReflectionE (:882-923): |(n2-1)/(1+n2)|, n2 = Re sqrt(eps)."""
    epsNum = V.plasmaFreqE * V.plasmaFreqE
    epsDom = (V.omega_0E * V.omega_0E - (2 * np.pi * P.freq_in * 2 * np.pi * P.freq_in)
              + 1j * V.gammaE * 2 * np.pi * P.freq_in)
    epsilon = 1 + epsNum / epsDom
    refr2 = np.real(np.sqrt(epsilon))
    return abs((refr2 - 1) / (1 + refr2))
