"""Drop-in for the reference's Environment_Setup.py (Environment_Setup.py:19-166).

envSetup derives the simulation parameters from the input frequency and
domain size and returns the 20-tuple consumed positionally by
MasterController.Params.  Reference behaviour kept: points per wavelength
(400, 350 for Kerr, 60*Re(eps)^1.05 when V exists), dt at Courant 0.95,
PML width 6*int(lambda/dz), the FFT-bin-aligned step search (whose
fallback leaves timeSteps = maxim-1, :101-116), the Nlam/200 extension and
every guard that prints and calls sys.exit().
"""
import sys

import numpy as np


def _exit(msg):
    print(msg)
    sys.exit()


def envSetup(newFreq_in, domainSize, minim=400, maxim=600, VExists=False, V=[], P=[], nonLinMed=False,
             LorMed=False):
    pi = np.pi
    epsilon = None
    if VExists:
        epsNum = V.plasmaFreqE * V.plasmaFreqE
        epsDom = V.omega_0E * V.omega_0E - (2 * pi * P.freq_in * 2 * pi * P.freq_in) + 1j * V.gammaE * 2 * pi * P.freq_in
        epsilon = 1 + epsNum / epsDom
    freq_in = newFreq_in
    c0 = 299792458.0
    lamMin = c0 / freq_in
    if VExists:
        Nlam = int(200 * (np.real(epsilon)) ** 1.05) if nonLinMed else int(60 * (np.real(epsilon)) ** 1.05)
    else:
        Nlam = 350 if nonLinMed else 400
    dz = lamMin / Nlam
    delT = (dz / c0) * 0.95
    period = 1 / freq_in
    courantNo = (c0 * delT) / dz
    if courantNo > 3 or courantNo < 0:
        _exit(f"{courantNo} courantNo is unstable")
    pmlWidth = 6 * int(lamMin / dz)
    if pmlWidth >= 12000:
        _exit(f"pmlWidth too big {pmlWidth}")
    Nz = int(domainSize / dz) + 2 * pmlWidth
    if minim == maxim:
        minim -= 1
    nonInt = False
    N = minim
    for N in range(minim, maxim):
        check = (freq_in * N) / (1 / delT)
        if int(check) - check == 0:
            break
        if N == maxim - 1:
            nonInt = True
    if nonInt:
        # the reference searches for the closest bin here but then keeps the
        # loop variable, i.e. maxim-1 (Environment_Setup.py:101-116)
        N = maxim - 1
    timeSteps = N if N >= minim else minim
    timeSteps += int(timeSteps * (Nlam / 200))
    if timeSteps >= 2 ** 15:
        _exit("timeSteps too large")
    nzsrcFromPml = int(0.05 / dz)
    if nzsrcFromPml >= Nz * 0.65:
        _exit(f"{nzsrcFromPml} src is too far into domain")
    nzsrc = nzsrcFromPml + pmlWidth
    if nzsrc - 10 <= pmlWidth:
        _exit("The probe for fft is in the PML region")
    MaterialDistFromPml = int(0.1 / dz)
    materialFrontEdge = MaterialDistFromPml + pmlWidth
    materialRearEdge = Nz - 1
    if materialRearEdge - materialFrontEdge < 10:
        _exit("width is too small or negative")
    if MaterialDistFromPml >= domainSize / dz:
        _exit("Material starts in CPML region")
    if materialFrontEdge <= nzsrc:
        _exit("Source is inside material")
    x1Loc = materialFrontEdge - 20
    x2Loc = nzsrc - 100
    eLoss = 0
    mLoss = 0
    return (Nz, timeSteps, eLoss, mLoss, (1 - eLoss) / (1 + eLoss), 1 / (1 + eLoss),
            (1 - mLoss) / (1 + mLoss), 1 / (1 + mLoss), x1Loc, x2Loc, materialFrontEdge,
            materialRearEdge, pmlWidth, nzsrc, lamMin, dz, delT, courantNo, period, Nlam)
