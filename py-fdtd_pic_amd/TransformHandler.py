"""Drop-in for the reference's TransformHandler.py output stage.

RefTester (TransformHandler.py:34-78) takes the FFT of a probe series and
returns (tail spectrum, tail frequencies, peak non-DC amplitude 2|Y|/T);
plots are not drawn.  The CZT the reference calls (MasterController.py:669)
does not exist there; `CZT` here evaluates the FP64 DFT of the series on a
frequency grid [f_lo, f_hi] with as many points as the series.
"""
import sys

import numpy as np


def RefTester(V, P, y, m, mul=False):
    y = np.asarray(y, dtype=np.float64)
    Y = np.fft.fft(y)
    Ypow = (2 * np.abs(Y)) / P.timeSteps
    freqs = np.fft.fftfreq(len(y), d=P.delT)
    indMax = int(np.argmax(Ypow))
    if indMax == 0:
        print("Could not find non-DC freq")
        sys.exit()
    val = Ypow[indMax]
    half = int(len(Y) / 2)
    bob = (np.abs(Y[half + 10:len(Y) - 500]) * m * 2) / len(Y)
    return bob, freqs[half + 10:len(freqs) - 500], val


def CZT(f_lo, f_hi, fs, y, t, n):
    y = np.asarray(y, dtype=np.float64)
    f = np.linspace(f_lo, f_hi, int(n))
    ph = np.exp(-2j * np.pi * np.outer(f, np.asarray(t, dtype=np.float64)))
    return f, ph @ y / fs
