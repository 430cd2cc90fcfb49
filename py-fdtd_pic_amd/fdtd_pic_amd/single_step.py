"""BaseFDTD11's per-step update functions, one device kernel each.

The fast path is the integrator (Solver_Engine): the whole step loop runs on
the device.  Code that calls the per-step functions one at a time, as the
reference's integrators do (Solver_Engine.py:297-316), gets each function
executed by the FP64 single-op kernels of libfdtdpic (fdtd_step_op, no FMA
contraction, the reference's expression order): the arrays go to the device
and back on every call, and the results equal the reference's bit for bit
(the Kerr root to the rounding of pow/acos/cos).  There is no host fallback:
without the library these functions raise.

Return values follow the reference (BaseFDTD11.py:609-877): the updated
array, or (psi, field) for the CPML updates; arrays are updated in place when
they are contiguous float64.
"""
import ctypes as C

import numpy as np

from . import _capi as capi
from . import engine

_ctx = None

OPS = {"ADE_ExUpdate": 0, "ADE_HyUpdate": 1, "CPML_Psi_e_Update": 2, "CPML_Psi_m_Update": 3,
       "ADE_DxUpdate": 4, "ADE_ExCreate": 5, "ADE_PolarisationCurrent_Ex": 6, "AcubicFinder": 7,
       "NonLinExUpdate": 8}


def context():
    global _ctx
    if _ctx is None:
        _ctx = engine.Context("f64")
    return _ctx


def run_op(op, arrays, scalars=(), ranges=((0, 0), (0, 0))):
    """Run op on host arrays (list of float64 arrays; the written ones first)."""
    ctx = context()
    bufs = [np.ascontiguousarray(a, dtype=np.float64) for a in arrays]
    ptrs = (capi._dp * len(bufs))(*[capi.dptr(b) for b in bufs])
    lens = (C.c_int64 * len(bufs))(*[b.size for b in bufs])
    sc = np.ascontiguousarray(scalars, dtype=np.float64) if len(scalars) else np.zeros(1)
    rg = (C.c_int32 * 4)(int(ranges[0][0]), int(ranges[0][1]), int(ranges[1][0]), int(ranges[1][1]))
    capi.check(ctx.lib.fdtd_step_op(ctx.h, int(op), len(bufs), ptrs, lens, len(scalars), capi.dptr(sc), rg))
    return bufs


def _store(obj, name, buf):
    cur = getattr(obj, name)
    if cur is not buf:
        setattr(obj, name, buf)
    return buf


def apply(name, V, P, C_V, counts=0):
    import BaseFDTD11
    op = OPS[name]
    N = P.Nz
    if name == "ADE_ExUpdate":
        b = run_op(op, [V.Ex, V.Hy, V.Jx, V.UpExMat, C_V.den_Exdz], (), ((1, N + 1), (0, 0)))
        return _store(V, "Ex", b[0])
    if name == "ADE_HyUpdate":
        b = run_op(op, [V.Hy, V.Ex, V.UpHySelf, V.UpHyMat, C_V.den_Hydz], (), ((1, N), (0, 0)))
        return _store(V, "Hy", b[0])
    if name in ("CPML_Psi_e_Update", "CPML_Psi_m_Update"):
        n, W = len(V.Ex), P.pmlWidth
        left = (1, W) if P.CPMLXm else (0, 0)
        if name == "CPML_Psi_e_Update":
            right = (n - W, n) if P.CPMLXp else (0, 0)
            b = run_op(op, [V.Ex, C_V.psi_Ex, V.Hy, C_V.beX, C_V.ceX, C_V.Cb], (), (left, right))
            return _store(C_V, "psi_Ex", b[1]), _store(V, "Ex", b[0])
        right = (n - W, n - 1) if P.CPMLXp else (0, 0)
        b = run_op(op, [V.Hy, C_V.psi_Hy, V.Ex, C_V.bmY, C_V.cmY, C_V.C2], (), (left, right))
        return _store(C_V, "psi_Hy", b[1]), _store(V, "Hy", b[0])
    F, R = P.materialFrontEdge, P.materialRearEdge
    if name == "ADE_DxUpdate":
        b = run_op(op, [V.Dx, V.Hy, C_V.den_Exdz], (P.delT / P.dz,), ((F, R), (0, 0)))
        return _store(V, "Dx", b[0])
    if name == "ADE_ExCreate":
        b = run_op(op, [V.Ex, V.Dx, V.polarisationCurr], (P.permit_0,), ((F, R), (0, 0)))
        return _store(V, "Ex", b[0])
    if name == "ADE_PolarisationCurrent_Ex":
        A, B, Cc = BaseFDTD11.ade_coefficients(V, P)
        b = run_op(op, [V.polarisationCurr, V.tempTempVarPol, V.Ex], (A, B, Cc), ((F, R), (0, 0)))
        return _store(V, "polarisationCurr", b[0])
    if name == "AcubicFinder":
        cub, qua, one = BaseFDTD11.cubic_coefficients(V, P)
        b = run_op(op, [V.Acubic, V.Dx], (cub, qua, one, P.permit_0), ((F, R), (0, 0)))
        return _store(V, "Acubic", b[0])
    if name == "NonLinExUpdate":
        b = run_op(op, [V.Ex, V.Dx, V.Acubic], (P.permit_0, V.chi3Stat, float(np.sqrt(1.2))), ((F, R), (0, 0)))
        return _store(V, "Ex", b[0])
    raise KeyError(name)
