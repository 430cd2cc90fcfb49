"""Known-answer pins for the oracle's small input/output functions, the last
ones checked only through runs that compose them:

  sig_mod (Solver_Engine.Sig_Mod / SechArr, Solver_Engine.py:126-140):
    (AmpMod + s) sech(2 pi t / tau) with the argument capped at 50.  The
    half-maximum of sech lies at acosh 2 = ln(2 + sqrt 3); the cap makes the
    tail constant; a constant input shows the offset.
  atten_probe (Solver_Engine.probeSim, attenRead, :31-38): probes every
    25 cells from the slab's front edge, fed E[i] = i so the recorded values
    are the indices themselves.
  vid_make (Solver_Engine.vidMake, :57-68): row int(counts / interval),
    rows past the history ignored.
  ref_tester (TransformHandler.RefTester, TransformHandler.py:34-65): the
    peak-bin amplitude 2|Y_k|/T of a cosine of amplitude A on bin k is A;
    a DC-only input has no non-DC peak.
  analytical_reflection (BaseFDTD11.AnalyticalReflectionE, :882-923): a
    Lorentz eps of exactly 4 gives |(2-1)/(2+1)| = 1/3; no plasma gives 0;
    eps < 0 (no real index) gives 1.
  physical.dft: X(f) = dt sum x_t exp(-2 pi i f t dt) of a cosine on a DFT
    bin is A T dt / 2; linear.
  gaussian_src (BaseFDTD11.Gaussian, :86-96): the peak 2A/Sc at step 440,
    1/e of it 200 steps either side, the exponent's alternating sign at f = 0.5.
  mat_setup (Material_Def.matSetup, Material_Def.py:21-120): the grid of a
    5 GHz call worked by hand: dz = c0/(200 f), delT = 0.99 dz/c0 =
    0.99e-12 s, 900 cells, source at 250, slab 450..749, probes at 350 and
    190, and 799 steps (no N in [400, 800) makes f N delT integral, so the
    nearest-search branch runs and leaves the loop variable at its last
    value: the reference's quirk).
Each test names the mistake it catches; test_planted_* plant one in the
oracle (monkeypatched, on the CPU) and show the pin rejects it.
"""
import math
from types import SimpleNamespace

import numpy as np
import pytest

from oracle import faithful as FO
from oracle import physical as ph


# ---------------------------------------------------------------------------
# sig_mod: sech envelope
# ---------------------------------------------------------------------------
def _sech_profile(fn, n=9000, dt=1e-12, tau=1e-9, amp=1.0, s=0.0):
    P = SimpleNamespace(delT=dt)
    return fn(P, np.full(n, s), AmpMod=amp, tau=tau)


def _sech_ok(out, dt=1e-12, tau=1e-9):
    t_half = math.log(2 + math.sqrt(3)) * tau / (2 * math.pi)      # acosh(2) / omega
    k = int(t_half / dt)                                           # 209 for these values
    return out[0] == 1.0 and out[k] > 0.5 > out[k + 1] and out[-1] == out[-2]


def test_sig_mod_half_maximum_cap_and_offset():
    out = _sech_profile(FO.sig_mod)
    assert _sech_ok(out)
    # the cap: omega t >= 50 from t = 50 tau / (2 pi) = 7.958 ns (index 7958 at 1 ps), then constant
    assert np.all(out[7958:] == out[-1]) and out[7956] > out[7957] > out[7958]
    # the offset multiplies the envelope: (AmpMod + s) at t = 0
    assert _sech_profile(FO.sig_mod, amp=1.0, s=0.5)[0] == 1.5


def test_planted_sig_mod_mistakes_fail():
    def tau_not_inverted(P, sig, AmpMod=1, tau=14.6e-10):         # omega = 2 pi tau
        a = np.minimum(2 * np.pi * tau * np.arange(len(sig)) * P.delT, 50)
        return (AmpMod + sig) / np.cosh(a)

    def no_cap(P, sig, AmpMod=1, tau=14.6e-10):                   # sech without the cap at 50
        a = 2 * np.pi / tau * np.arange(len(sig)) * P.delT
        with np.errstate(over="ignore"):
            return (AmpMod + sig) / np.cosh(a)

    def cos_not_cosh(P, sig, AmpMod=1, tau=14.6e-10):
        a = np.minimum(2 * np.pi / tau * np.arange(len(sig)) * P.delT, 50)
        return (AmpMod + sig) / np.cos(a)
    for bad in (tau_not_inverted, no_cap, cos_not_cosh):
        assert not _sech_ok(_sech_profile(bad)), bad.__name__


# ---------------------------------------------------------------------------
# atten_probe, vid_make: index rules
# ---------------------------------------------------------------------------
def _atten(fn):
    V = SimpleNamespace(Ex=np.arange(400, dtype=float), attenAmnt=3, x1Atten=np.zeros((3, 10)))
    P = SimpleNamespace(materialFrontEdge=100)
    fn(V, P, 7)
    return V.x1Atten


def test_atten_probe_records_every_25th_cell_from_the_front():
    got = _atten(FO.atten_probe)
    assert np.array_equal(got[:, 7], [100.0, 125.0, 150.0])
    assert not got[:, :7].any() and not got[:, 8:].any()


def test_planted_atten_probe_stride_fails():
    def stride_from_zero(V, P, counts, granAttn=25):
        inds = np.arange(0, granAttn * V.attenAmnt, granAttn)
        for i in range(len(V.x1Atten)):
            V.x1Atten[i][counts] = V.Ex[inds[i]]
    assert not np.array_equal(_atten(stride_from_zero)[:, 7], [100.0, 125.0, 150.0])


def test_vid_make_row_rule():
    V = SimpleNamespace(Ex_History=np.zeros((4, 5)), Ex=None)
    P = SimpleNamespace(vidInterval=10)
    for counts in (0, 20, 35, 45):
        V.Ex = np.full(5, float(counts))
        FO.vid_make(V, P, counts)
    # rows 0, 2, 3 (int(3.5) = 3); 45 -> row 4 lies past the history and is ignored
    assert np.array_equal(V.Ex_History[:, 0], [0.0, 0.0, 20.0, 35.0])


# ---------------------------------------------------------------------------
# ref_tester: peak-bin amplitude
# ---------------------------------------------------------------------------
def test_ref_tester_amplitude_of_a_bin_cosine():
    T = 1000
    P = SimpleNamespace(timeSteps=T)
    y = 0.8 * np.cos(2 * np.pi * 37 * np.arange(T) / T + 0.3)
    assert abs(FO.ref_tester(P, y) - 0.8) < 1e-12
    with pytest.raises(FO.RefExit):
        FO.ref_tester(P, np.full(T, 2.0))


# ---------------------------------------------------------------------------
# analytical_reflection: normal-incidence Fresnel of the Lorentz eps
# ---------------------------------------------------------------------------
def _lorentz(wp2_over_w2, w0_over_w, gamma=0.0, f=1e14):
    w = 2 * np.pi * f
    V = SimpleNamespace(plasmaFreqE=math.sqrt(wp2_over_w2) * w, omega_0E=w0_over_w * w, gammaE=gamma)
    return V, SimpleNamespace(freq_in=f)


def test_analytical_reflection_known_answers():
    # eps = 1 + wp^2/(w0^2 - w^2) = 1 + 9/3 = 4 -> n = 2 -> R = 1/3
    assert abs(FO.analytical_reflection(*_lorentz(9.0, 2.0)) - 1 / 3) < 1e-12
    assert FO.analytical_reflection(*_lorentz(0.0, 2.0)) == 0.0
    # above resonance with a strong plasma term: eps = 1 + 10/(0.25 - 1) < 0, no real index
    assert abs(FO.analytical_reflection(*_lorentz(10.0, 0.5)) - 1.0) < 1e-12


def test_planted_reflection_without_the_vacuum_term_fails():
    def no_one(V, P):                                    # eps = wp^2 / (...) without the 1
        w = 2 * np.pi * P.freq_in
        eps = V.plasmaFreqE ** 2 / (V.omega_0E ** 2 - w * w + 1j * V.gammaE * w)
        n = np.real(np.sqrt(eps))
        return abs((n - 1) / (1 + n))
    assert abs(no_one(*_lorentz(9.0, 2.0)) - 1 / 3) > 1e-3


# ---------------------------------------------------------------------------
# physical.dft
# ---------------------------------------------------------------------------
def test_dft_of_a_bin_cosine_and_linearity():
    T, dt = 512, 2e-15
    f0 = 40 / (T * dt)
    t = np.arange(T) * dt
    x = 1.7 * np.cos(2 * np.pi * f0 * t)
    X = ph.dft(x[None, :], dt, np.array([f0]))[0, 0]
    assert abs(X - 1.7 * T * dt / 2) < 1e-12 * T * dt
    y = np.sin(2 * np.pi * 3 * f0 * t)
    fr = np.array([f0, 2.5 * f0])
    assert np.allclose(ph.dft((2 * x + y)[None, :], dt, fr), 2 * ph.dft(x[None, :], dt, fr)
                       + ph.dft(y[None, :], dt, fr), rtol=0, atol=1e-12 * T * dt)


# ---------------------------------------------------------------------------
# mat_setup: the 5 GHz grid worked by hand
# ---------------------------------------------------------------------------
def test_mat_setup_hand_worked_5ghz():
    r = FO.mat_setup(None, None, None, None, 5e9, 600)
    (Nz, T, eL, mL, ca, cb, da, db, x1, x2, front, rear, W, src, lam, dz, dt, cn, period) = r
    c0 = 299792458.0
    assert (Nz, T, x1, x2, front, rear, W, src) == (900, 799, 350, 190, 450, 749, 150, 250)
    assert (eL, mL, ca, cb, da, db) == (0, 0, 1.0, 1.0, 1.0, 1.0)
    assert abs(dz - c0 / (200 * 5e9)) < 1e-18 and abs(dt - 0.99e-12) < 1e-25
    assert abs(cn - 0.99) < 1e-15 and abs(period - 2e-10) < 1e-25 and abs(lam - c0 / 5e9 * 0.1) < 1e-15


# ---------------------------------------------------------------------------
# gaussian_src (BaseFDTD11.Gaussian, :86-96): 2 A exp(-(t - 440)^2 / 200^2
# cos(2 pi f (t - 440))) / Sc, t in steps: the cosine sits inside the
# exponent (the reference's form), so with f = 1 it is a plain Gaussian and
# with f = 0.5 the exponent's sign alternates step by step
# ---------------------------------------------------------------------------
def _gauss(fn, f):
    P = SimpleNamespace(timeSteps=1000, freq_in=f, Amplitude=1.0, courantNo=0.5)
    return np.asarray(fn(P))


def _gauss_ok(g1, gh):
    e = math.exp(1.0)
    return (abs(g1[440] - 4.0) < 1e-15 and abs(g1[640] - 4.0 / e) < 1e-14 and abs(g1[240] - 4.0 / e) < 1e-14
            and abs(gh[441] - 4.0 * math.exp(1.0 / 40000.0)) < 1e-14
            and abs(gh[442] - 4.0 * math.exp(-4.0 / 40000.0)) < 1e-14)


def test_gaussian_source_known_values():
    assert _gauss_ok(_gauss(FO.gaussian_src, 1.0), _gauss(FO.gaussian_src, 0.5))


def test_planted_modulated_gaussian_fails():
    def textbook(P):                                   # cos as a carrier outside the exponent
        t = np.arange(P.timeSteps)
        env = np.exp(-((t - 440.0) ** 2) / 40000.0)
        return P.Amplitude * 2 * env * np.cos(2 * np.pi * P.freq_in * (t - 440)) / P.courantNo
    assert not _gauss_ok(_gauss(textbook, 1.0), _gauss(textbook, 0.5))
