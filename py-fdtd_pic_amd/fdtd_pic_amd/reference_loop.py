"""Map the reference's parameter objects onto one device simulation.

The reference integrators (Solver_Engine.py:142-371) run, per step:
ExUpdate over n = 1..N, CPML psi_E over [1,W) U [N+1-W, N], source/TFSF,
Dx/ExCreate (or the Kerr E) over the slab, HyUpdate over 1..N-1 and CPML
psi_H over [1,W) U [N+1-W, N-1].  `reference_spec` turns the arrays the
reference set-up builds (V.UpExMat/UpHyMat, C_V.den_*, beX/ceX/Cb,
bmY/cmY/C2) into the per-cell coefficients of fdtd_sim_desc_t, so the device
executes the same arithmetic; `run_pass` advances it and copies the results
back into V / C_V.
"""
import numpy as np

from . import _capi as capi
from .engine import Context, SimSpec


def _psi_masks(P, n):
    W = P.pmlWidth
    me = np.zeros(n, dtype=bool)
    mh = np.zeros(n, dtype=bool)
    if P.CPMLXm:
        me[1:W] = True
        mh[1:W] = True
    if P.CPMLXp:
        me[n - W:n] = True          # CPML_Psi_e_Update: range(len-W, len)  (:373)
        mh[n - W:n - 1] = True      # CPML_Psi_m_Update: range(len-W, len-1) (:390)
    return me, mh


def reference_spec(V, P, C_V, mode, ade, srcE, srcH, probes, snap=None):
    """SimSpec for one pass of a reference integrator.

    mode: "linlor" (D/ExCreate slab), "nl" (D + Kerr cubic slab), "free" (no D).
    srcE/srcH: the per-step values the loop adds, Exs[t]/Sc and Hys[t]/Sc.
    probes: list of (cell, t0, t1).  snap: (interval, rows) or None.
    """
    N, T = int(P.Nz), int(P.timeSteps)
    n = N + 1
    if np.any(np.asarray(V.UpHySelf[1:N]) != 1.0):
        raise NotImplementedError("UpHySelf != 1 is not supported by the device loop")
    ke = np.asarray(V.UpExMat, dtype=np.float64) * np.asarray(C_V.den_Exdz, dtype=np.float64)
    kh = np.asarray(V.UpHyMat, dtype=np.float64) * np.asarray(C_V.den_Hydz, dtype=np.float64)
    ke[0] = 0.0                      # ExUpdate runs n = 1..N (:665)
    kh[0] = 0.0
    kh[N] = 0.0                      # HyUpdate runs n = 1..N-1 (:643)
    me, mh = _psi_masks(P, n)
    z = np.zeros(n)
    bE = np.where(me, C_V.beX, z)
    cE = np.where(me, C_V.ceX, z)
    CbE = np.where(me, C_V.Cb, z)
    bH = np.where(mh, C_V.bmY, z)
    cH = np.where(mh, C_V.cmY, z)
    C2H = np.where(mh, C_V.C2, z)
    F, R = int(P.materialFrontEdge), int(P.materialRearEdge)
    mid = n // 2
    ke0 = float(np.asarray(V.UpExMat)[mid] * np.asarray(C_V.den_Exdz)[mid])
    kh0 = float(np.asarray(V.UpHyMat)[mid] * np.asarray(C_V.den_Hydz)[mid])
    s = SimSpec(N, T)
    if mode == "free":
        # background values of UpdateCoef (BaseFDTD11.py:169-176); the slab divides by eps/mu
        ke0 = float(P.CharImp * P.courantNo)
        kh0 = float((1 / P.CharImp) * P.courantNo)
        keS, khS = float(ke[F] if F < n else ke0), float(kh[F] if F < n else kh0)
        if keS != ke0 or khS != kh0:
            s.F, s.R = F, R
            s.slab_kind = capi.SLAB_DIEL
            s.keS, s.khS = keS, khS
            s.set_cells(ke, kh, bE, cE, CbE, bH, cH, C2H, ke0, kh0, diel=(F, R, keS, khS))
        else:
            s.set_cells(ke, kh, bE, cE, CbE, bH, cH, C2H, ke0, kh0)
    else:
        s.set_cells(ke, kh, bE, cE, CbE, bH, cH, C2H, ke0, kh0)
        den = np.asarray(C_V.den_Exdz, dtype=np.float64)[F:R]
        if np.any(den != 1.0):
            raise NotImplementedError("kappa != 1 inside the slab is not supported by the device loop")
        s.F, s.R = F, R
        s.kD = P.delT / P.dz
        s.jD = 0.0
        ns = R - F
        if mode == "linlor":
            s.slab_kind = capi.SLAB_LINEAR
            s.sl_e = np.full(ns, P.permit_0)
        else:
            from BaseFDTD11 import cubic_coefficients
            s.slab_kind = capi.SLAB_KERR_REF
            s.kerr_a, s.kerr_b, s.kerr_c = cubic_coefficients(V, P)
            s.kerr_e0 = P.permit_0
            s.kerr_lin = float(np.real(np.sqrt(1.2)))
            s.kerr_chi3 = float(V.chi3Stat)
            s.sl_e = np.ones(ns)
        if ade:
            from BaseFDTD11 import ade_coefficients
            A, B, Cc = ade_coefficients(V, P)
            s.ade = 1
            s.sl_A, s.sl_B, s.sl_C = np.full(ns, A), np.full(ns, B), np.full(ns, Cc)
    s.src_cell = int(P.nzsrc)
    s.srcE = np.asarray(srcE, dtype=np.float64)
    if P.TFSF:
        s.tfsf = 1
        s.srcH = np.asarray(srcH, dtype=np.float64)
    s.probes = list(probes)
    if snap is not None:
        s.snap_interval, s.snap_rows = int(snap[0]), int(snap[1])
    return s


def read_pass(sim, spec):
    """The arrays a pass leaves behind (fields, psi, slab D/P/P^{n-1}, probes,
    snapshots) as a dict of host arrays."""
    out = {k: sim.read(k) for k in ("E", "H", "PSIE", "PSIH")}
    if spec.R > spec.F and spec.slab_kind >= capi.SLAB_LINEAR:
        for k in ("D", "P", "PM"):
            out[k] = sim.read(k)
    out["probes"] = sim.probes()
    if spec.snap_rows:
        out["snap"] = sim.snapshots()
    return out


def run_passes(specs, precision="f64"):
    """Run passes together on the device (one context); returns their dicts."""
    with Context(precision) as ctx:
        sims = [ctx.add(spec) for spec in specs]
        ctx.run_all()
        return [read_pass(sim, spec) for sim, spec in zip(sims, specs)]


def run_pass(spec, precision="f64"):
    """Run one pass on the device; returns a dict of host arrays."""
    return run_passes([spec], precision)[0]


def write_back(V, C_V, P, spec, out):
    """Copy a pass's device results into the reference objects."""
    V.Ex = out["E"]
    V.Hy = out["H"]
    C_V.psi_Ex = out["PSIE"]
    C_V.psi_Hy = out["PSIH"]
    if "D" in out:
        F, R = spec.F, spec.R
        V.Dx = np.zeros(spec.N + 1)
        V.Dx[F:R] = out["D"]
        V.polarisationCurr = np.zeros(spec.N + 1)
        V.polarisationCurr[F:R] = out["P"]
        V.tempVarPol = np.zeros(spec.N + 1)
        V.tempVarPol[F:R] = out["PM"]
    if spec.snap_rows:
        vi = spec.snap_interval
        rows = [t // vi for t in range(vi, spec.steps, vi) if t // vi < spec.snap_rows]
        for r in rows:
            V.Ex_History[r] = out["snap"][r]
