"""Oracle: FP64 physical model for the configs the reference cannot run.  TEST INFRASTRUCTURE ONLY.

The reference loop (Solver_Engine.py:294-316) is restated here with the
physics the workloads ask for: eps_inf != 1, per-cell Lorentz parameters, an
E-form Kerr root, a textbook CPML with the correct sign, a soft
Gaussian-modulated source and an optional current sheet J from the PIC
deposit.  Step order is the reference's: ADE polarisation (:299), E update
(:301), CPML psi_E (:305), source (:307), D update + E creation (:312-313),
H update (:314), CPML psi_H (:316), probes (:360-370).

Units are normalised so that every Yee coefficient is the Courant number:
    E [V/m],  H' = Z0*H,  D' = D/eps0,  P' = P/eps0,  J' = Z0*Jx_ref
where Jx_ref = J_x*dz is the reference's Jx (BaseFDTD11.py:667).
    vacuum:  E  += Sc*(H'[n]-H'[n-1] - J'[n])
    slab:    P'n = A P' + B P'm + C E ;  D' += Sc*(dH' - J') ;
             E   = root of chi3 E^3 + eps_inf E = D' - P'n
    H:       H'  += Sc*(E[n+1]-E[n])
    CPML:    psi = b psi + c*diff ; field += Sc*psi   (kappa = 1, alpha = 0)
A, B, C are BaseFDTD11.py:620-626 with eps0*wp^2 -> wp^2 (wp^2 = d_eps*w0^2).

A spec is a plain dict; `build` turns it into coefficient arrays.
"""
import math

import numpy as np

from .cubic import kerr_root_e

C0 = 299792458.0


def _cellwise(v, n):
    a = np.asarray(v, dtype=np.float64)
    return np.full(n, float(a)) if a.ndim == 0 else a.astype(np.float64).copy()


def default_spec(N=2000, steps=10000, **kw):
    """Config-1 shape: lambda0 = 1.5 um, dz = lambda0/40, Sc = 0.5, 400-cell slab."""
    spec = dict(
        N=N, steps=steps, Sc=0.5, dz=1.5e-6 / 40,
        pml_width=32, pml_order=3, pml_R0=1e-8,
        slab_start=N // 2, slab_stop=N // 2 + 400,
        eps_inf=1.5, d_eps=2.0, omega0=2 * math.pi * 150e12, gamma=2 * math.pi * 5e12,
        chi3=0.0, ade=True,
        src_cell=100, src_f0=200e12, src_tau=30e-15, src_t0=120e-15, src_amp=1.0,
        probes=[60, N - 32 - 10],
    )
    spec.update(kw)
    return spec


def pml_profile(spec):
    """Per-cell CPML b, c for E nodes (z = n dz) and H nodes (z = (n+1/2) dz).

    sigma*dt/eps0 = Sc * (m+1) * (-ln R0) / (2 W) * depth^m, alpha = 0, kappa = 1,
    so b = exp(-sigma dt/eps0) and c = b - 1.
    """
    N, W, m = spec["N"], spec["pml_width"], spec["pml_order"]
    n = np.arange(N + 1, dtype=np.float64)
    sig0 = spec["Sc"] * (m + 1) * (-math.log(spec["pml_R0"])) / (2.0 * W)
    dE = np.zeros(N + 1)
    dH = np.zeros(N + 1)
    left = n < W
    dE[left] = (W - n[left]) / W
    dH[left] = (W - n[left] - 0.5) / W
    right = n >= N - W
    dE[right] = (n[right] - (N - W)) / W
    dH[right] = np.minimum((n[right] + 0.5 - (N - W)) / W, 1.0)
    bE = np.exp(-sig0 * dE ** m)
    bH = np.exp(-sig0 * dH ** m)
    return bE, bE - 1.0, bH, bH - 1.0, dE > 0, dH > 0


def ade_coefficients(spec, dt, omega0, d_eps):
    """A, B, C of the Lorentz ADE (BaseFDTD11.py:620-626) in normalised P' = P/eps0."""
    g = spec["gamma"]
    D = 1.0 / dt ** 2 + g / (2.0 * dt)
    A = (2.0 / dt ** 2 - omega0 ** 2) / D
    B = (g / (2.0 * dt) - 1.0 / dt ** 2) / D
    C = (d_eps * omega0 ** 2) / D
    return A, np.full_like(A, B), C


def source_table(spec):
    """Soft source s_t = amp exp(-((t dt - t0)/tau)^2) sin(2 pi f0 (t dt - t0))."""
    dt = spec["Sc"] * spec["dz"] / C0
    t = np.arange(spec["steps"], dtype=np.float64) * dt - spec["src_t0"]
    return spec["src_amp"] * np.exp(-(t / spec["src_tau"]) ** 2) * np.sin(2 * math.pi * spec["src_f0"] * t)


def build(spec, with_slab=True):
    N = spec["N"]
    dt = spec["Sc"] * spec["dz"] / C0
    F, R = spec["slab_start"], spec["slab_stop"]
    ns = R - F
    m = dict(N=N, Sc=float(spec["Sc"]), dt=dt, steps=int(spec["steps"]))
    m["bE"], m["cE"], m["bH"], m["cH"], m["pmlE"], m["pmlH"] = pml_profile(spec)
    m["F"], m["R"] = (F, R) if with_slab else (F, F)
    eps = _cellwise(spec["eps_inf"], ns)
    w0 = _cellwise(spec["omega0"], ns)
    de = _cellwise(spec["d_eps"], ns)
    if spec.get("ade", True):
        A, B, C = ade_coefficients(spec, dt, w0, de)
    else:
        A, B, C = np.zeros(ns), np.zeros(ns), np.zeros(ns)
    m["A"], m["B"], m["C"], m["eps"] = A, B, C, eps
    m["chi3"] = float(spec["chi3"])
    m["src_cell"] = int(spec["src_cell"])
    m["src"] = source_table(spec)
    m["probes"] = [int(p) for p in spec["probes"]]
    return m


class State:
    def __init__(self, N, ns):
        self.E = np.zeros(N + 1)
        self.H = np.zeros(N + 1)
        self.psiE = np.zeros(N + 1)
        self.psiH = np.zeros(N + 1)
        self.D = np.zeros(ns)
        self.P = np.zeros(ns)
        self.Pm = np.zeros(ns)


def kerr_vec(u, eps, chi3):
    if chi3 == 0.0:
        return u / eps
    return np.array([kerr_root_e(ui, ei, chi3) for ui, ei in zip(u, eps)])


def step(m, s, t, J=None):
    """One full step in the reference order; J is J' (normalised), length N+1."""
    N, Sc = m["N"], m["Sc"]
    F, R = m["F"], m["R"]
    E, H = s.E, s.H
    dH = np.zeros(N + 1)
    dH[1:N] = H[1:N] - H[0:N - 1]
    Jv = np.zeros(N + 1) if J is None else J
    # slab polarisation uses E at the start of the step
    if R > F:
        Pn = m["A"] * s.P + m["B"] * s.Pm + m["C"] * E[F:R]
        s.Pm = s.P
        s.P = Pn
    # E update n = 1..N-1 (E[0], E[N] PEC)
    Enew = E.copy()
    Enew[1:N] = E[1:N] + Sc * (dH[1:N] - Jv[1:N])
    pe = m["pmlE"].copy()
    pe[0] = pe[N] = False
    s.psiE[pe] = m["bE"][pe] * s.psiE[pe] + m["cE"][pe] * dH[pe]
    Enew[pe] += Sc * s.psiE[pe]
    Enew[m["src_cell"]] += m["src"][t]
    if R > F:
        s.D = s.D + Sc * (dH[F:R] - Jv[F:R])
        Enew[F:R] = kerr_vec(s.D - s.P, m["eps"], m["chi3"])
    s.E = E = Enew
    # H update n = 0..N-1
    dE = E[1:N + 1] - E[0:N]
    s.H[0:N] = H[0:N] + Sc * dE
    ph = m["pmlH"].copy()
    ph[N] = False
    idx = np.nonzero(ph)[0]
    s.psiH[idx] = m["bH"][idx] * s.psiH[idx] + m["cH"][idx] * dE[idx]
    s.H[idx] += Sc * s.psiH[idx]


def run(m, steps=None, record=True):
    steps = m["steps"] if steps is None else steps
    s = State(m["N"], m["R"] - m["F"])
    probes = np.zeros((len(m["probes"]), steps))
    for t in range(steps):
        step(m, s, t)
        if record:
            probes[:, t] = s.E[m["probes"]]
    return s, probes


def dft(x, dt, freqs):
    """FP64 DFT X(f) = sum_t x_t exp(-2 pi i f t dt) dt."""
    t = np.arange(x.shape[-1]) * dt
    ph = np.exp(-2j * math.pi * np.outer(freqs, t))
    return (ph @ x.T).T * dt


def spectra(spec, freqs=None):
    """Two-pass protocol (Solver_Engine.py:276): pass 0 vacuum, pass 1 with slab.

    R(f) = |X1_r - X0_r| / |X0_r| at probe 0;  T(f) = |X1_t| / |X0_t| at probe 1.
    """
    if freqs is None:
        freqs = np.linspace(100e12, 300e12, 256)
    m0 = build(spec, with_slab=False)
    m1 = build(spec, with_slab=True)
    _, p0 = run(m0)
    s1, p1 = run(m1)
    X0 = dft(p0, m0["dt"], freqs)
    X1 = dft(p1, m1["dt"], freqs)
    R = np.abs(X1[0] - X0[0]) / np.abs(X0[0])
    T = np.abs(X1[1]) / np.abs(X0[1])
    return freqs, R, T, s1, p0, p1
