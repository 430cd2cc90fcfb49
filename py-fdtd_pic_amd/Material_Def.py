"""Drop-in for the reference's Material_Def.py (Material_Def.py:21-120).

matSetup is the older parameter builder: fixed Nlam = 20 on lambda/10,
dt at Courant 0.99, PML 150, domain 600 cells, and a 19-tuple (no Nlam).
"""
import sys

import numpy as np


def _exit(msg):
    print(msg)
    sys.exit()


def matSetup(V, P, C_P, C_V, newFreq_in, domainSize, minim=400, maxim=800):
    freq_in = newFreq_in
    c0 = 299792458.0
    lamMin = (c0 / freq_in) * 0.1
    Nlam = 20
    dz = lamMin / Nlam
    delT = 0.99 / (c0 * (1 / dz))
    period = 1 / freq_in
    courantNo = (c0 * delT) / dz
    if courantNo > 1 or courantNo < 0:
        _exit(f"{courantNo} courantNo is unstable")
    pmlWidth = 150
    domainSize = 600
    Nz = domainSize + 2 * pmlWidth
    N = minim
    for N in range(minim, maxim):
        check = (freq_in * N) / (1 / delT)
        if int(check) - check == 0:
            break
    timeSteps = N
    if timeSteps >= 10000:
        _exit("timeSteps too large")
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
