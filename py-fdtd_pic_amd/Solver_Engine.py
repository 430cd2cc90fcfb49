"""Drop-in for the reference's Solver_Engine.py.

Same entry points, signatures and 8-tuple returns as the reference
(Solver_Engine.py:142-371).  Each integrator pass does the reference's host
set-up (FieldInit, EmptySpaceCalc, CPML builders, spatialStab, sources), then
runs the whole `for counts in range(P.timeSteps)` loop as one device
simulation (fdtd_pic_amd.reference_loop -> libfdtdpic.so) and copies the
fields, CPML psi, probes and snapshots back into V and C_V.
"""
import numpy as np

import BaseFDTD11
import genericStability as gStab
from fdtd_pic_amd import reference_loop as RL

PRECISION = "f64"


def probeSim(V, P, C_V, C_P, counts, val="null", af=False, granAttn=25, whichField="Ex",
             attenRead=False, nonlinear=False):
    """Solver_Engine.probeSim (:16-54): one probe sample (host-side bookkeeping)."""
    if not nonlinear:
        if attenRead:
            inds = np.arange(P.materialFrontEdge, granAttn * V.attenAmnt + P.materialFrontEdge, granAttn)
            for i in range(len(V.x1Atten)):
                V.x1Atten[i][counts] = V.Ex[inds[i]]
            return V.x1Atten
        if af:
            V.x1ColAf[counts] = val
            return V.x1ColAf
        V.x1ColBe[counts] = val
        return V.x1ColBe
    V.Port1[counts] = V.Ex[P.materialFrontEdge]
    V.Port2[counts] = V.Ex[P.materialRearEdge]
    return V.Port1, V.Port2


def vidMake(V, P, C_V, C_P, counts, field, whichField="Ex"):
    """Solver_Engine.vidMake (:57-68)."""
    shrunkCnt = int(counts / P.vidInterval)
    if whichField == "Ex":
        if shrunkCnt < len(V.Ex_History):
            V.Ex_History[shrunkCnt] = field
        return V.Ex_History
    return "vidMake has gone to the end of the returns "


def boundCondManager(V, P, C_V, C_P):
    """Solver_Engine.boundCondManager (:70-87)."""
    if P.CPMLXp or P.CPMLXm:
        BaseFDTD11.CPML_ScalingCalc(V, P, C_V, C_P)
        BaseFDTD11.CPML_Ex_RC_Define(V, P, C_V, C_P)
        BaseFDTD11.CPML_HY_RC_Define(V, P, C_V, C_P)
        BaseFDTD11.CPML_Ex_Update_Coef(V, P, C_V, C_P)
        BaseFDTD11.CPML_Hy_Update_Coef(V, P, C_V, C_P)
        BaseFDTD11.denominators(V, P, C_V, C_P)
    return C_V


def SourceManager(V, P, C_V, C_P):
    """Solver_Engine.SourceManager (:89-124): SineCont (with the NL pump) or Gaussian tables."""
    if P.SineCont:
        Exs, Hys1 = BaseFDTD11.SmoothTurnOn(V, P)
        Exp, Hyp1 = np.zeros(P.timeSteps), np.zeros(P.timeSteps)
        if P.nonLinMed:
            Exp, Hyp1 = BaseFDTD11.SmoothTurnOn(V, P, tempfreq=P.freq_in * 0.8)
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
        Exs = BaseFDTD11.Gaussian(V, P)
        Hys = BaseFDTD11.Gaussian(V, P) if P.TFSF else np.zeros(len(Exs))
        return Exs, Hys
    return [], []


def SechArr(a):
    a = np.where(a > 50, 50, a)
    return 1 / np.cosh(a)


def Sig_Mod(V, P, sig, AmpCarr=1, AmpMod=1, tau=14.6e-10):
    """Solver_Engine.Sig_Mod (:133-140): sech envelope."""
    trang = np.arange(len(sig)) * P.delT
    omega_mod = 2 * np.pi * (1 / tau)
    return (AmpMod + sig) * SechArr(omega_mod * trang)


def _per_step(P, Exs, Hys):
    """Values the loop adds each step: Exs[t]/Sc and Hys[t]/Sc (:307-310)."""
    Exs = np.asarray(Exs, dtype=np.float64)
    Hys = np.asarray(Hys, dtype=np.float64) if len(Hys) else np.zeros(P.timeSteps)
    if len(Exs) < P.timeSteps:
        raise IndexError("source table shorter than timeSteps (no source selected?)")
    return Exs[:P.timeSteps] / P.courantNo, Hys[:P.timeSteps] / P.courantNo


def _atten_probes(V, P, t0, t1, granAttn=25):
    inds = np.arange(P.materialFrontEdge, granAttn * V.attenAmnt + P.materialFrontEdge, granAttn)
    return [(int(c), t0, t1) for c in inds[:len(V.x1Atten)]]


def _store_probes(V, P, i, pr, finish, start):
    T = P.timeSteps
    if i == 0:
        hi = min(finish, T - 1) + 1
        V.x1ColBe[:hi] = pr[0, :hi]
    else:
        V.x1ColAf[start:T] = pr[0, start:T]
        if P.atten:
            for k in range(len(V.x1Atten)):
                V.x1Atten[k][start:T] = pr[1 + k, start:T]


class PassPlan:
    """The device passes of one integrator call, planned but not run.

    The host set-up of every pass (FieldInit, coefficients, CPML, spatialStab,
    sources) runs when the plan is made, in the reference's order, so plans
    made one after the other leave V / P exactly as the sequential calls
    would.  `passes` holds (SimSpec, post) pairs: post(out) copies one pass's
    device results into V / C_V.  `result()` is the integrator's 8-tuple once
    every post has run.  `execute` runs the passes of many plans as one batch
    on the device."""

    def __init__(self):
        self.passes = []
        self.result = None


def plan_free_space(V, P, C_V, C_P, probeReadFinishBe, probeReadStartAf):
    """The two passes of IntegratorFreeSpace1D (:142-215), julia=False."""
    plan = PassPlan()
    T = P.timeSteps
    Exs = Hys = None
    for i in range(0, 2):
        BaseFDTD11.FieldInit(V, P)
        V.UpHyMat, V.UpExMat = BaseFDTD11.EmptySpaceCalc(V, P)
        BaseFDTD11.Material(V, P)
        V.UpHyMat, V.UpExMat = BaseFDTD11.UpdateCoef(V, P)
        C_V = BaseFDTD11.CPML_FieldInit(V, P, C_V, C_P)
        C_V = boundCondManager(V, P, C_V, C_P)
        Exs, Hys = SourceManager(V, P, C_V, C_P)
        tauIn = 1 / (P.freq_in / 5)
        Exs = Sig_Mod(V, P, Exs, tau=tauIn)
        Hys = Sig_Mod(V, P, Hys, AmpMod=1 / P.CharImp, tau=tauIn)
        sE, sH = _per_step(P, Exs, Hys)
        if i == 0:
            probes = [(P.x1Loc, 0, min(probeReadFinishBe, T - 1))]
        else:
            probes = [(P.x2Loc, probeReadStartAf, T - 1)]
            if P.atten:
                probes += _atten_probes(V, P, probeReadStartAf, T - 1)
        snap = (P.vidInterval, len(V.Ex_History)) if i == 1 else None
        spec = RL.reference_spec(V, P, C_V, "free", False, sE, sH, probes, snap)
        plan.passes.append((spec, _pass_post(V, P, C_V, spec, i, probeReadFinishBe, probeReadStartAf)))
    plan.result = lambda: (V.Ex, V.Hy, Exs, Hys, C_V.psi_Ex, C_V.psi_Hy, V.x1ColBe, V.x1ColAf)
    return plan


def _pass_post(V, P, C_V, spec, i, finish, start):
    def post(out):
        RL.write_back(V, C_V, P, spec, out)
        _store_probes(V, P, i, out["probes"], finish, start)
    return post


def plan_nl(V, P, C_V, C_P, probeReadFinishBe, probeReadStartAf):
    """The single pass of IntegratorNL1D (:220-271): Kerr cubic on the slab, ports at F and R."""
    plan = PassPlan()
    T = P.timeSteps
    BaseFDTD11.FieldInit(V, P)
    V.UpHyMat, V.UpExMat = BaseFDTD11.EmptySpaceCalc(V, P)
    C_V = BaseFDTD11.CPML_FieldInit(V, P, C_V, C_P)
    C_V = boundCondManager(V, P, C_V, C_P)
    _, _, _, V.plasmaFreqE, _ = gStab.spatialStab(P.timeSteps, P.Nz, P.dz, P.freq_in, P.delT,
                                                   V.plasmaFreqE, V.omega_0E, V.gammaE)
    Exs, Hys = SourceManager(V, P, C_V, C_P)
    sE, sH = _per_step(P, Exs, Hys)
    probes = [(P.materialFrontEdge, 0, T - 1), (P.materialRearEdge, 0, T - 1)]
    spec = RL.reference_spec(V, P, C_V, "nl", False, sE, sH, probes, (P.vidInterval, len(V.Ex_History)))

    def post(out):
        RL.write_back(V, C_V, P, spec, out)
        V.Port1 = np.array(out["probes"][0])
        V.Port2 = np.array(out["probes"][1])
    plan.passes.append((spec, post))
    plan.result = lambda: (V.Ex, V.Hy, Exs, Hys, C_V.psi_Ex, C_V.psi_Hy, V.x1ColBe, V.x1ColAf)
    return plan


def plan_linlor(V, P, C_V, C_P, probeReadFinishBe, probeReadStartAf):
    """The two passes of IntegratorLinLor1D (:275-371): pass 0 incident, pass 1 Lorentz slab."""
    plan = PassPlan()
    T = P.timeSteps
    Exs = Hys = None
    for i in range(0, 2):
        BaseFDTD11.FieldInit(V, P)
        V.UpHyMat, V.UpExMat = BaseFDTD11.EmptySpaceCalc(V, P)
        C_V = BaseFDTD11.CPML_FieldInit(V, P, C_V, C_P)
        C_V = boundCondManager(V, P, C_V, C_P)
        _, _, _, V.plasmaFreqE, _ = gStab.spatialStab(P.timeSteps, P.Nz, P.dz, P.freq_in, P.delT,
                                                       V.plasmaFreqE, V.omega_0E, V.gammaE)
        Exs, Hys = SourceManager(V, P, C_V, C_P)
        V.test = 0
        sE, sH = _per_step(P, Exs, Hys)
        if i == 0:
            probes = [(P.x1Loc, 0, min(probeReadFinishBe, T - 1))]
        else:
            probes = [(P.x2Loc, probeReadStartAf, T - 1)]
            if P.atten:
                probes += _atten_probes(V, P, probeReadStartAf, T - 1)
        snap = (P.vidInterval, len(V.Ex_History)) if i == 1 else None
        spec = RL.reference_spec(V, P, C_V, "linlor", i == 1, sE, sH, probes, snap)
        plan.passes.append((spec, _pass_post(V, P, C_V, spec, i, probeReadFinishBe, probeReadStartAf)))
    plan.result = lambda: (V.Ex, V.Hy, Exs, Hys, C_V.psi_Ex, C_V.psi_Hy, V.x1ColBe, V.x1ColAf)
    return plan


def execute(plans, precision=PRECISION, scheduler=None):
    """Run every pass of the plans on the device, then apply the posts in plan
    order.  Without a scheduler the passes share one context (they all run
    together); with one (fdtd_pic_amd.sweep.Scheduler) each plan is a job
    admitted under its HBM budget.  Returns each plan's result()."""
    if scheduler is None:
        outs = RL.run_passes([spec for pl in plans for spec, _ in pl.passes], precision)
        k = 0
        for pl in plans:
            for _, post in pl.passes:
                post(outs[k])
                k += 1
    else:
        from fdtd_pic_amd import sweep
        jobs = [sweep.SpecJob(i, [spec for spec, _ in pl.passes]) for i, pl in enumerate(plans)]
        scheduler.run_batch(jobs)
        for pl, job in zip(plans, jobs):
            for (_, post), out in zip(pl.passes, job.outputs):
                post(out)
    return [pl.result() for pl in plans]


def IntegratorFreeSpace1D(V, P, C_V, C_P, probeReadFinishBe, probeReadStartAf):
    """Solver_Engine.IntegratorFreeSpace1D (:142-215), julia=False: both passes on the device."""
    return execute([plan_free_space(V, P, C_V, C_P, probeReadFinishBe, probeReadStartAf)])[0]


def IntegratorNL1D(V, P, C_V, C_P, probeReadFinishBe, probeReadStartAf):
    """Solver_Engine.IntegratorNL1D (:220-271)."""
    return execute([plan_nl(V, P, C_V, C_P, probeReadFinishBe, probeReadStartAf)])[0]


def IntegratorLinLor1D(V, P, C_V, C_P, probeReadFinishBe, probeReadStartAf):
    """Solver_Engine.IntegratorLinLor1D (:275-371): pass 0 incident, pass 1 Lorentz slab, on the device together."""
    return execute([plan_linlor(V, P, C_V, C_P, probeReadFinishBe, probeReadStartAf)])[0]
