"""Oracle: the per-step update functions of the reference's BaseFDTD.py, as
plain Python loops over cells.  TEST INFRASTRUCTURE ONLY.

Each function restates the BaseFDTD.py function it names, one cell at a
time in the reference's loop order and expression order, on copies of the
arrays, and returns the arrays the reference updates.  The product runs
these on the device (py-fdtd_pic_amd/BaseFDTD.py, fdtd_step_op ops 1, 5 and
9-13); tests/test_gpu_single_step.py compares the two bit for bit and
checks the updates against closed forms (a frozen-difference psi series, a
skipped slab, Ex = (Dx - P) / eps0).
"""
import numpy as np


def _c(a):
    return np.array(a, dtype=np.float64, copy=True)


def cpml_psi_e_update(Ex, psi, Hy, beX, ceX, Cb, Nz):
    """BaseFDTD.py:331-336."""
    Ex, psi = _c(Ex), _c(psi)
    for nz in range(1, Nz - 1):
        psi[nz] = beX[nz] * psi[nz] + ceX[nz] * (Hy[nz] - Hy[nz - 1])
        Ex[nz] = Ex[nz] + Cb[nz] * psi[nz]
    return psi, Ex


def cpml_psi_m_update(Hy, psi, Ex, bmY, cmY, C2, Nz):
    """BaseFDTD.py:339-344."""
    Hy, psi = _c(Hy), _c(psi)
    for nz in range(0, Nz - 2):
        psi[nz] = bmY[nz] * psi[nz] + cmY[nz] * (Ex[nz] - Ex[nz + 1])
        Hy[nz] = Hy[nz] + C2[nz] * psi[nz]
    return psi, Hy


def cpml_hy_update(Hy, Ex, UpHySelf, UpHyMat, UpHyEcompsCo, Nz):
    """BaseFDTD.py:349-352."""
    Hy = _c(Hy)
    for nz in range(0, Nz - 2):
        Hy[nz] = Hy[nz] * UpHySelf[nz] + (Ex[nz + 1] - Ex[nz]) * UpHyMat[nz] * UpHyEcompsCo[nz]
    return Hy


def cpml_ex_update(Ex, Hy, UpExSelf, UpExHcompsCo, UpExMat, Nz):
    """BaseFDTD.py:355-358."""
    Ex = _c(Ex)
    for nz in range(1, Nz - 1):
        Ex[nz] = Ex[nz] * UpExSelf[nz] + (Hy[nz] - Hy[nz - 1]) * UpExHcompsCo[nz] * UpExMat[nz]
    return Ex


def ade_polarisation_current_ex(pol, tvp, ttvp, Ex, gammaE, omega_0E, plasmaFreqE, delT, permit_0, front, rear):
    """BaseFDTD.py:378-390."""
    pol = _c(pol)
    D = (1 / delT ** 2) + (gammaE / (2 * delT))
    A = ((2 / delT ** 2) - omega_0E ** 2) / D
    B = ((gammaE / (2 * delT)) - 1 / delT ** 2) / D
    C = (permit_0 * (plasmaFreqE ** 2)) / D
    for nz in range(int(front - 1), int(rear)):
        pol[nz] = A * tvp[nz] + B * ttvp[nz] + C * Ex[nz]
    return pol


def ade_hy_update(Hy, Ex, UpHySelf, UpHyMat, Nz):
    """BaseFDTD.py:392-395."""
    Hy = _c(Hy)
    for nz in range(1, Nz - 1):
        Hy[nz] = Hy[nz] * UpHySelf[nz] + (Ex[nz + 1] - Ex[nz]) * UpHyMat[nz]
    return Hy


def ade_ex_update(Ex, Hy, UpExSelf, UpExMat, front, rear, Nz):
    """BaseFDTD.py:401-408."""
    Ex = _c(Ex)
    for nz in range(1, int(front - 1)):
        Ex[nz] = UpExSelf[nz] * Ex[nz] + (Hy[nz] - Hy[nz - 1]) * UpExMat[nz]
    if rear < Nz - 1:
        for nzz in range(int(rear - 1), Nz):
            Ex[nzz] = UpExSelf[nzz] * Ex[nzz] + (Hy[nzz] - Hy[nzz - 1]) * UpExMat[nzz]
    return Ex


def ade_ex_create(Ex, Dx, pol, permit_0, front, rear):
    """BaseFDTD.py:411-414."""
    Ex = _c(Ex)
    for nz in range(int(front - 1), int(rear)):
        Ex[nz] = (Dx[nz] - pol[nz]) / permit_0
    return Ex


def ade_dx_update(Dx, Ex, Hy, pol, delT, dz, permit_0, front, rear):
    """BaseFDTD.py:417-422 (a loop from cell 0 reads Hy[-1], the last cell)."""
    Dx, Ex = _c(Dx), _c(Ex)
    for nz in range(int(front - 1), int(rear)):
        Dx[nz] = Dx[nz] + (Hy[nz] - Hy[nz - 1]) * delT / dz
        Ex[nz] = (Dx[nz] - pol[nz]) / permit_0
    return Dx, Ex
