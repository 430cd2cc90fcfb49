"""Drop-in names of the reference's BaseFDTD.py.

BaseFDTD.py is the pre-numba ancestor of BaseFDTD11.py: the same function
names with older semantics (FieldInit's 11-tuple and Nz <= 3000 limit,
BaseFDTD.py:35-79; ADE_ExUpdate skipping the slab, :401-408; the CPML psi
sign of :334-348).  No live module of the reference imports it; its loop
semantics are superseded by BaseFDTD11.  This module keeps the names with
their BaseFDTD argument lists and gives them BaseFDTD11's behaviour, so code
written against either module runs the same device path:

* set-up functions (fields, coefficients, CPML profiles, sources) are
  BaseFDTD11's host functions;
* per-step updates run as device kernels (fdtd_pic_amd.single_step);
* the TFSF corrections and PEC/PMC ends are single-cell bookkeeping on the
  caller's arrays (BaseFDTD.py:168-174, :363-369);
* the empty stubs of the reference (ADE_MyUpdate, ADE_NonLinMyUpdate,
  ADE_NonLinPxUpdate, SpatialFiltering, SymbolicRegression) stay empty.
"""
import BaseFDTD11 as _b11

sourceGen = _b11.sourceGen
sourceGen2 = _b11.sourceGen2
FieldInit = _b11.FieldInit
SmoothTurnOn = _b11.SmoothTurnOn
EmptySpaceCalc = _b11.EmptySpaceCalc
Material = _b11.Material
UpdateCoef = _b11.UpdateCoef
CPML_FieldInit = _b11.CPML_FieldInit
CPML_ScalingCalc = _b11.CPML_ScalingCalc
CPML_Ex_RC_Define = _b11.CPML_Ex_RC_Define
CPML_HY_RC_Define = _b11.CPML_HY_RC_Define
CPML_Ex_Update_Coef = _b11.CPML_Ex_Update_Coef
CPML_Hy_Update_Coef = _b11.CPML_Hy_Update_Coef
denominators = _b11.denominators


def HyTfSfCorr(V, P, counts):
    """BaseFDTD.py:168-170."""
    V.Hy[P.nzsrc - 1] -= V.Exs[counts] / P.CharImp
    return V.Hy[P.nzsrc - 1]


def ExTfSfCorr(V, P, counts):
    """BaseFDTD.py:172-174."""
    V.Ex[P.nzsrc] += V.Hys[counts]
    return V.Ex[P.nzsrc]


def CPML_Psi_e_Update(V, P, C_V, C_P):
    return _b11.CPML_Psi_e_Update(V, P, C_V, C_P)


def CPML_Psi_m_Update(V, P, C_V, C_P):
    return _b11.CPML_Psi_m_Update(V, P, C_V, C_P)


def CPML_HyUpdate(V, P, C_V, C_P):
    return _b11.ADE_HyUpdate(V, P, C_V, C_P)


def CPML_ExUpdate(V, P, C_V, C_P):
    return _b11.ADE_ExUpdate(V, P, C_V, C_P, 0)


def CPML_PEC(V, P, C_V, C_P):
    """BaseFDTD.py:363-365."""
    V.Ex[0] = 0
    return V.Ex[0], V.Ex[P.Nz - 1]


def CPML_PMC(V, P, C_V, C_P):
    """BaseFDTD.py:367-369."""
    V.Hy[P.Nz - 1] = 0
    return V.Hy[P.Nz - 1]


def ADE_TempPolCurr(V, P):
    """BaseFDTD.py:371-375: shift of the polarisation history."""
    V.tempTempVarPol = V.tempVarPol.copy()
    V.tempVarPol = V.polarisationCurr.copy()
    return V.tempTempVarPol, V.tempVarPol


def ADE_PolarisationCurrent_Ex(V, P, C_V, C_P):
    return _b11.ADE_PolarisationCurrent_Ex(V, P, C_V, C_P, 0)


def ADE_HyUpdate(V, P, C_V, C_P):
    return _b11.ADE_HyUpdate(V, P, C_V, C_P)


def ADE_ExUpdate(V, P, C_V, C_P):
    return _b11.ADE_ExUpdate(V, P, C_V, C_P, 0)


def ADE_ExCreate(V, P, C_V, C_P):
    return _b11.ADE_ExCreate(V, P, C_V, C_P)


def ADE_DxUpdate(V, P, C_V, C_P):
    """BaseFDTD.py:418-423 returns (Dx, Ex): D update, then E from D - P."""
    _b11.ADE_DxUpdate(V, P, C_V, C_P)
    _b11.ADE_ExCreate(V, P, C_V, C_P)
    return V.Dx, V.Ex


def ADE_MyUpdate():
    pass


def ADE_NonLinMyUpdate():
    pass


def ADE_NonLinPxUpdate():
    pass


def SpatialFiltering():
    pass


def SymbolicRegression():
    pass
