"""Drop-in for the reference's MasterController.py: parameter objects and Controller.

The reference defines four numba jitclasses (MasterController.py:82-440)
whose member types round what is stored: Params.courantNo is float32
(:234,247), Variables.alpha3 is float32 (:129), CPML_Params.kappaMax /
r_scale / r_a_scale / alphaMax are float32 (:341-347), duplicate spec keys
resolve last-wins.  The classes here are plain Python with the same member
names and the same rounding on assignment, so every coefficient the
integrators derive from them is the reference's value.

Unlike the reference, importing this module runs nothing (the reference
runs __Main__ at import, :683); call __Main__() explicitly.
"""
import math
import time as tim

import numpy as np

import Environment_Setup as envDef
import Solver_Engine as SE
import TransformHandler as transH
import BaseFDTD11

EPS0 = 8.8541878188e-12    # scipy.constants.epsilon_0, CODATA 2022 (MasterController.py:291)
MU0 = 1.25663706127e-06    # scipy.constants.mu_0 (:292)


def _f32(v):
    return float(np.float32(v))


def _i32(v):
    return int(np.int32(int(v)))


class _Typed:
    """Rounds assignments the way the reference's jitclass spec does."""
    _spec = {}

    def __setattr__(self, name, value):
        kind = self._spec.get(name)
        if kind == "f32":
            value = _f32(value)
        elif kind == "f64":
            value = float(value)
        elif kind == "i32":
            value = _i32(value)
        elif kind == "bool":
            value = bool(value)
        elif kind == "arr":
            value = np.asarray(value, dtype=np.float64)
        object.__setattr__(self, name, value)


# effective (last-wins) member types of specP (MasterController.py:217-284)
_SPEC_P = dict(
    Nz="i32", timeSteps="i32", eLoss="f64", mLoss="f64", eSelfCo="f64", eHcompsCo="f64",
    hSelfCo="f64", hEcompsCo="f64", x1Loc="i32", x2Loc="i32", materialFrontEdge="i32",
    materialRearEdge="i32", pmlWidth="i32", nzsrc="i32", lamMin="f64", dz="f64", delT="f64",
    courantNo="f32", period="f64", domainSize="i32", freq_in="f64", permit_0="f64",
    permea_0="f64", CharImp="f64", c0="f64", Nlam="f64", MORmode="bool", delayMOR="i32",
    CPMLXp="bool", CPMLXm="bool", TFSF="bool", SineCont="bool", Gaussian="bool", Ricker="bool",
    Amplitude="f32", Periods="f32", LorentzMed="bool", nonLinMed="bool", FreeSpace="bool",
    epsRe="f32", muRe="f32", vidMake="bool", vidInterval="i32", atten="bool", julia="bool",
    testMode="bool")


class Params(_Typed):
    """MasterController.Params (MasterController.py:287-339)."""
    _spec = _SPEC_P

    def __init__(self, Nz, timeSteps, eLoss, mLoss, eSelfCo, eHcompsCo, hSelfCo, hEcompsCo, x1Loc,
                 x2Loc, materialFrontEdge, materialRearEdge, pmlWidth, nzsrc, lamMin, dz, delT,
                 courantNo, period, Nlam, MORmode, domainSize, freq_in, delayMOR, LorentzMed=False,
                 nonLinMed=False, SineCont=False, Gaussian=False, TFSF=False):
        self.permit_0 = EPS0
        self.permea_0 = MU0
        self.CharImp = 376.730313668
        self.c0 = 299792458.0
        self.freq_in = freq_in
        self.lamMin = lamMin
        self.Nlam = Nlam
        self.dz = dz
        self.delT = delT
        self.courantNo = courantNo
        self.materialFrontEdge = materialFrontEdge
        self.materialRearEdge = materialRearEdge
        self.Nz = Nz
        self.timeSteps = timeSteps
        self.x1Loc = x1Loc
        self.x2Loc = x2Loc
        self.nzsrc = nzsrc
        self.period = period
        self.eLoss = eLoss
        self.eSelfCo = eSelfCo
        self.eHcompsCo = eHcompsCo
        self.mLoss = mLoss
        self.hSelfCo = hSelfCo
        self.hEcompsCo = hEcompsCo
        self.pmlWidth = pmlWidth
        self.domainSize = domainSize
        self.MORmode = MORmode
        self.delayMOR = delayMOR
        self.CPMLXp = True
        self.CPMLXm = True
        self.TFSF = TFSF
        self.SineCont = SineCont
        self.Gaussian = Gaussian
        self.Ricker = False
        self.Amplitude = 1.0
        self.Periods = 1.0
        self.LorentzMed = LorentzMed
        self.nonLinMed = nonLinMed
        self.FreeSpace = True
        self.epsRe = 1.0
        self.muRe = 1.0
        self.vidMake = True
        self.vidInterval = 50
        self.atten = False
        self.julia = False
        self.testMode = False

    def __str__(self):
        return 'Class containing all values that remain constant throughout a sim'


_SPEC_V = dict(Nz="i32", timeSteps="f32", plasmaFreqE="f64", gammaE="f64", omega_0E="f64",
               test="i32", tempTempTest="i32", attenAmnt="i32", alpha3="f32",
               nonLin3gammaE="f64", nonLin3Omega_0E="f64", chi1Stat="f64", chi3Stat="f64")


class Variables(_Typed):
    """MasterController.Variables (MasterController.py:145-213)."""
    _spec = _SPEC_V

    def __init__(self, Nz, timeSteps, interval, attenAmnt):
        n = Nz + 1
        for name in ("UpHyMat", "UpExMat", "UpExHcompsCo", "UpExSelf", "UpHySelf", "UpHyEcompsCo",
                     "epsilon", "mu"):
            setattr(self, name, np.ones(n))
        for name in ("Ex", "Hy", "Dx", "Jx", "polarisationCurr", "tempVarPol", "tempTempVarPol",
                     "tempVarE", "tempTempVarE", "tempVarHy", "tempTempVarHy", "tempVarJx",
                     "tempTempVarJx", "tempVarDx", "tempTempVarDx", "tempTest", "Gx3", "Qx3",
                     "Pbar3", "JxKerr", "JxRaman", "Acubic"):
            setattr(self, name, np.zeros(n))
        rows = int(timeSteps / interval)
        for name in ("Ex_History", "Hy_History", "Jx_History", "polCurr_History", "Dx_History",
                     "Psi_e_History"):
            setattr(self, name, np.zeros((rows, n)))
        for name in ("x1ColBe", "x1ColAf", "x1Jx", "x1Hy", "x1ExOld", "x1JxOld", "x1HyOld",
                     "Port1", "Port2"):
            setattr(self, name, np.zeros(timeSteps))
        self.plasmaFreqE = np.sqrt((1.5) * (2 * np.pi * 20e9) ** 2)
        self.gammaE = 2 * np.pi * 20e9 * 0.1
        self.omega_0E = 2 * np.pi * 20e9
        self.test = 5
        self.tempTempTest = 0
        self.attenAmnt = attenAmnt
        self.x1Atten = np.zeros((attenAmnt, timeSteps))
        self.alpha3 = 0.7
        self.nonLin3gammaE = 0
        self.nonLin3Omega_0E = 6e9
        self.chi1Stat = np.sqrt(1.2) - 1
        self.chi3Stat = 1e-3
        self.cubPoly = np.zeros((n, 4), dtype=np.complex128)
        self.roots = np.zeros(4, dtype=np.complex128)

    def __str__(self):
        return 'Contains data that will change during sim'


class CPML_Params(_Typed):
    """MasterController.CPML_Params (MasterController.py:341-366)."""
    _spec = dict(kappaMax="f32", r_scale="f32", r_a_scale="f32", sigmaEMax="f64", sigmaHMax="f64",
                 sigmaOpt="f64", alphaMax="f32")

    def __init__(self, dz):
        self.kappaMax = 1
        self.r_scale = 4
        self.r_a_scale = 1
        self.sigmaEMax = 0.5 * (0.8 * (1) / (dz * (MU0 / EPS0) ** 0.5))
        self.sigmaHMax = self.sigmaEMax
        self.sigmaOpt = self.sigmaEMax
        self.alphaMax = 0.05

    def __repr__(self):
        return f'{self.__class__.__name__}'

    def __str__(self):
        return 'Class containing all CPML values that remain constant throughout a sim'


class CPML_Variables(_Typed):
    """MasterController.CPML_Variables (MasterController.py:401-440)."""
    _spec = dict(Nz="i32")

    def __init__(self, Nz, timeSteps):
        n = Nz + 1
        for name in ("kappa_Ex", "kappa_Hy", "psi_Ex", "psi_Hy", "alpha_Ex", "alpha_Hy", "sigma_Ex",
                     "sigma_Hy", "beX", "bmY", "ceX", "cmY", "Ca", "Cb", "Cc", "C1", "C2", "C3",
                     "eLoss_CPML", "mLoss_CPML", "den_Hydz", "den_Exdz", "tempTempVarPsiEx",
                     "tempVarPsiEx", "tempTempVarPsiHy", "tempVarPsiHy"):
            setattr(self, name, np.zeros(n))
        for name in ("psi_Ex_Probe", "psi_Hy_Probe", "psi_Ex_Old", "psi_Hy_Old"):
            setattr(self, name, np.zeros(timeSteps))

    def __repr__(self):
        return f'{self.__class__.__name__}'

    def __str__(self):
        return 'Class containing all CPML values that vary throughout a sim'


def Controller(V, P, C_V, C_P):
    """MasterController.Controller (MasterController.py:451-469): integrator dispatch
    LorentzMed > FreeSpace > nonLinMed, same probe windows."""
    probeReadFinishBe = int(P.timeSteps * 0.7)
    probeReadStartAf = int(P.timeSteps * 0.05)
    V.x1ColBe = np.zeros(P.timeSteps)
    V.x1ColAf = np.zeros(P.timeSteps)
    Exs = Hys = None
    if P.LorentzMed:
        V.Ex, V.Hy, Exs, Hys, C_V.psi_Ex, C_V.psi_Hy, V.x1ColBe, V.x1ColAf = SE.IntegratorLinLor1D(
            V, P, C_V, C_P, probeReadFinishBe, probeReadStartAf)
    elif P.FreeSpace:
        V.Ex, V.Hy, Exs, Hys, C_V.psi_Ex, C_V.psi_Hy, V.x1ColBe, V.x1ColAf = SE.IntegratorFreeSpace1D(
            V, P, C_V, C_P, probeReadFinishBe, probeReadStartAf)
    elif P.nonLinMed:
        V.Ex, V.Hy, Exs, Hys, C_V.psi_Ex, C_V.psi_Hy, V.x1ColBe, V.x1ColAf = SE.IntegratorNL1D(
            V, P, C_V, C_P, probeReadFinishBe, probeReadStartAf)
    return V, P, C_V, C_P, Exs, Hys


def results(V, P, C_V, C_P, time_Vec, RefCo=False, FFT=False, AnalRefCo=False, attenRead=False):
    """MasterController.results (MasterController.py:472-505), without plotting."""
    if attenRead:
        val = transH.RefTester(V, P, V.x1ColBe, 1)[2]
        return np.array([transH.RefTester(V, P, V.x1Atten[i], 1, mul=True)[2] / val
                         for i in range(len(V.x1Atten))])
    if RefCo:
        val = transH.RefTester(V, P, V.x1ColBe, 1)[2]
        val2 = transH.RefTester(V, P, V.x1ColAf, 1)[2]
        return val2 / val
    if AnalRefCo:
        return BaseFDTD11.AnalyticalReflectionE(V, P)
    return "results ran to end"


def _plan_controller(V, P, C_V, C_P):
    """Controller's dispatch (MasterController.py:451-469) as a plan: the host
    set-up of its integrator runs now, the device passes later (Solver_Engine.execute)."""
    probeReadFinishBe = int(P.timeSteps * 0.7)
    probeReadStartAf = int(P.timeSteps * 0.05)
    V.x1ColBe = np.zeros(P.timeSteps)
    V.x1ColAf = np.zeros(P.timeSteps)
    if P.LorentzMed:
        return SE.plan_linlor(V, P, C_V, C_P, probeReadFinishBe, probeReadStartAf)
    if P.FreeSpace:
        return SE.plan_free_space(V, P, C_V, C_P, probeReadFinishBe, probeReadStartAf)
    if P.nonLinMed:
        return SE.plan_nl(V, P, C_V, C_P, probeReadFinishBe, probeReadStartAf)
    return None


def LoopedSim(Rep, V, P, C_V, C_P, MORmode, domainSize, lowLimTim, highLimTim,
              stringparamSweep="Input frequency sweep", loop=False, Low=3e9, Interval=1e8,
              RefCoBool=True, points=20, batched=True):
    """MasterController.LoopedSim (MasterController.py:530-606) without plots/video.

    loop=True runs the "Input frequency sweep" and returns
    (freqs, measured reflection, analytical reflection, per-point
    (V, P, C_V, C_P)) in addition to the last (V, P, C_V, C_P, Exs, Hys).  The reference runs the points one after
    another (:543-559).  With batched=True (the default) the host set-up of
    every point runs first, in the reference's order (each point's envSetup
    reads the previous point's V and P, as there), then all points' passes
    run on the device together as jobs of one fdtd_pic_amd.sweep.Scheduler
    (FP64, admitted under its HBM budget), and each point's results are
    taken from its own V / P.  The analytical reflection is evaluated when
    its point is planned, at that point's freq_in, as the reference does
    before it steps freq_in on.  batched=False runs them one at a time.
    """
    Exs = Hys = None
    if loop:
        dataRange = np.arange(Low, points * Interval + Low, Interval)
        ref = np.zeros(points)
        anal = np.zeros(points)
        if stringparamSweep != "Input frequency sweep":
            return V, P, C_V, C_P, Exs, Hys, (dataRange, ref, anal, [])
        plans, objs = [], []
        for k in range(points):
            setup = envDef.envSetup(P.freq_in, domainSize, lowLimTim, highLimTim, VExists=True, V=V, P=P)
            P = Params(*setup, P.MORmode, domainSize, P.freq_in, 20, LorentzMed=P.LorentzMed,
                       SineCont=P.SineCont, Gaussian=P.Gaussian, TFSF=P.TFSF)
            V = Variables(P.Nz, P.timeSteps, P.vidInterval, 1)
            C_P = CPML_Params(P.dz)
            C_V = CPML_Variables(P.Nz, P.timeSteps)
            if batched:
                plan = _plan_controller(V, P, C_V, C_P)
                plans.append(plan)
                objs.append((V, P, C_V, C_P))
                anal[k] = results(V, P, C_V, C_P, None, AnalRefCo=True)
            else:
                V, P, C_V, C_P, Exs, Hys = Controller(V, P, C_V, C_P)
                t = np.arange(0, len(V.x1ColBe)) * P.delT
                ref[k] = results(V, P, C_V, C_P, t, RefCo=True)
                anal[k] = results(V, P, C_V, C_P, t, AnalRefCo=True)
                objs.append((V, P, C_V, C_P))
            P.freq_in = P.freq_in + Interval
        if batched:
            from fdtd_pic_amd import sweep
            sch = sweep.Scheduler(SE.PRECISION)
            try:
                outs = SE.execute([pl for pl in plans if pl is not None], SE.PRECISION, scheduler=sch)
            finally:
                sch.close()
            it = iter(outs)
            for k, (pl, (Vk, Pk, Ck, CPk)) in enumerate(zip(plans, objs)):
                if pl is not None:
                    Exs, Hys = next(it)[2:4]
                t = np.arange(0, len(Vk.x1ColBe)) * Pk.delT
                ref[k] = results(Vk, Pk, Ck, CPk, t, RefCo=True)
            V, P, C_V, C_P = objs[-1]
        return V, P, C_V, C_P, Exs, Hys, (dataRange, ref, anal, objs)
    tic = tim.perf_counter()
    V, P, C_V, C_P, Exs, Hys = Controller(V, P, C_V, C_P)
    if Rep is not None:
        Rep.printer(str(tim.perf_counter() - tic), "ControllerTime")
    return V, P, C_V, C_P, Exs, Hys


class Reporter:
    """MasterController.Reporter (MasterController.py:57-80)."""

    def __init__(self):
        self.dict1 = {"": ""}

    def printer(self, item="", name="", show=False):
        self.dict1[name] = item
        if show:
            print(item)


def __Main__(freq_in=9e9, domainSize=0.7, lowLimTim=7000, highLimTim=8000, nonLinMed=True, LorMed=False):
    """MasterController.__Main__ (MasterController.py:620-681) minus plots/video/CZT."""
    setupReturn = envDef.envSetup(freq_in, domainSize, lowLimTim, highLimTim, nonLinMed=nonLinMed,
                                  LorMed=LorMed)
    P = Params(*setupReturn, False, domainSize, freq_in, 20)
    P.vidInterval = 50
    V = Variables(P.Nz, P.timeSteps, P.vidInterval, 10)
    C_P = CPML_Params(P.dz)
    C_V = CPML_Variables(P.Nz, P.timeSteps)
    Rep = Reporter()
    P.atten = False
    P.epsRe = 1
    P.CPMLXp = True
    P.CPMLXm = True
    P.TFSF = True
    P.Gaussian = False
    P.SineCont = True
    P.Periods = 1000
    P.LorentzMed = LorMed
    P.nonLinMed = nonLinMed
    if P.nonLinMed:
        P.FreeSpace = False
        P.LorentzMed = False
    if P.LorentzMed:
        P.FreeSpace = False
        P.nonLinMed = False
    V, P, C_V, C_P, Exs, Hys = LoopedSim(Rep, V, P, C_V, C_P, False, domainSize, lowLimTim, highLimTim)
    return V, P, C_P, C_V, Rep
