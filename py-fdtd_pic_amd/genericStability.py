"""Drop-in for the reference's genericStability.spatialStab (genericStability.py:12-62).

Lorentz numerical-dispersion check.  It returns the plasma frequency scaled
by sqrt(|fix|); the integrators assign it back to V.plasmaFreqE on every
pass, so the rescale is cumulative (Solver_Engine.py:231,286).
"""
import sys

import numpy as np

C0 = 299792458.0


def spatialStab(fullTime, fullSpace, spaceStep, frequency, timeStep, plasmaFreq, resonantFreq, gamma,
                lim_Of_Stability=np.pi / 2):
    sp = spaceStep
    rad = 2 * np.pi * frequency
    twoPi = rad / frequency
    tim = timeStep
    if rad * tim > np.pi / 2 and frequency <= 5e9:
        print("unstable timestep", rad * tim, tim)
        sys.exit()
    sw = np.sinc(np.pi * frequency * tim)      # numpy sinc: sin(pi x)/(pi x)
    pf = plasmaFreq
    rf = resonantFreq
    es = (pf ** 2) / (rf ** 2) - 1
    gam = gamma
    sqN = rad ** 2 * sw ** 2 - es * rf ** 2 * np.cos(rad * tim) + 1.0j * gam * rad * sw
    sqD = rad ** 2 * sw ** 2 - rf ** 2 * np.cos(rad * tim) + 1.0j * gam * rad * sw
    arg = (rad / C0) * (sp / 2) * sw * np.sqrt(sqN / sqD)
    kNum = abs((2 / sp) * np.arcsin(arg))
    epsilon = 1 + (pf ** 2) / (rf ** 2 - rad ** 2 - 1j * gam * rad)
    refr = np.sqrt(abs(np.real(epsilon)))
    fix = (C0 * tim * np.sin((kNum * refr * sp) / 2)) / (refr * sp * np.sin((kNum * C0 * tim) / 2))
    pf = np.sqrt(abs(fix)) * pf
    lamCont = C0 / frequency
    lamDisc = twoPi / kNum
    diff = abs(lamCont - lamDisc)
    if kNum * sp > np.pi / 2:
        print("unstable, wave is not resolved:", kNum * sp)
        sys.exit()
    if kNum * sp * fullSpace < (5 * lamDisc):
        print("unstable because domain too small,", (lamDisc * 5) / sp)
        sys.exit()
    return lamCont, lamDisc, diff, pf, fix
