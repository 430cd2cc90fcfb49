"""Pins for oracle/physical.py (and its C restatement oracle/fdtd_cpu.c).

Nothing in the reference pins these numerically (its two unit tests do not
run: test_BaseFDTD.py:14-21 unpacks 6 of FieldInit's 11 returns with an
undefined Mc, test_sandbox.py:10-11 raises NameError), so each part is
checked against what the mathematics fixes:
  Yee vacuum update   - exact translation at Sc = 1 (magic time step);
                        numerical dispersion sin(w dt/2) = Sc sin(k dz/2);
                        discrete energy conservation in a closed cavity.
  D-form slab         - lossless dielectric slab: |R|^2 + |T|^2 = 1 and
                        Fabry-Perot minima at k_s L = m pi with the medium's
                        numerical dispersion sin(w dt/2) = (Sc/n) sin(k dz/2).
  Lorentz ADE         - second-order convergence of the discrete susceptibility
                        to Delta_eps w0^2/(w0^2 - w^2 - i gamma w); absorption
                        makes |R|^2 + |T|^2 < 1.
  Kerr                - the final slab state satisfies the constitutive law
                        chi3 E^3 + eps_inf E = D' - P'.
  CPML                - reflection below 1e-5 of the incident pulse.
  C restatement       - bitwise equal to the NumPy oracle.
"""
import math

import numpy as np
import pytest

from oracle import physical as ph

C0 = 299792458.0


def _vacuum_model(N, steps, Sc=0.5, pml=True, **kw):
    spec = ph.default_spec(N=N, steps=steps, Sc=Sc, slab_start=N // 2, slab_stop=N // 2, **kw)
    m = ph.build(spec)
    if not pml:
        m["pmlE"][:] = False
        m["pmlH"][:] = False
    return spec, m


def test_magic_time_step_translation():
    spec, m = _vacuum_model(600, 500, Sc=1.0, pml=False, src_cell=100, probes=[200, 201, 260],
                            src_t0=40e-15, src_tau=8e-15)
    _, pr = ph.run(m)
    a, b, c = pr
    # right-going wave: E[c+1](t+1) == E[c](t), E[c+60](t+60) == E[c](t)
    assert np.max(np.abs(pr)) > 0.1
    assert np.max(np.abs(b[1:] - a[:-1])) <= 1e-14 * np.max(np.abs(a))
    assert np.max(np.abs(c[60:] - a[:-60])) <= 1e-13 * np.max(np.abs(a))


def test_vacuum_numerical_dispersion():
    spec, m = _vacuum_model(3000, 8000, Sc=0.5, probes=[400, 1400])
    _, pr = ph.run(m)
    dt = m["dt"]
    dz = spec["dz"]
    L = 1000 * dz
    freqs = np.linspace(185e12, 215e12, 7)
    X = ph.dft(pr, dt, freqs)
    ratio = X[1] / X[0]
    assert np.all(np.abs(np.abs(ratio) - 1) < 1e-6)        # lossless propagation
    w = 2 * math.pi * freqs
    k_num = (2 / dz) * np.arcsin(np.sin(w * dt / 2) / spec["Sc"])
    phase = -np.angle(ratio)
    # compare modulo 2 pi
    d = (phase - k_num * L + math.pi) % (2 * math.pi) - math.pi
    assert np.max(np.abs(d)) < 1e-7
    # the continuum k = w/c is measurably different: the check is sharp
    dc = (phase - w / C0 * L + math.pi) % (2 * math.pi) - math.pi
    assert np.max(np.abs(dc)) > 1e-2


def test_energy_conserved_in_closed_cavity():
    spec, m = _vacuum_model(400, 1, Sc=0.5, pml=False, src_cell=150, src_t0=20e-15, src_tau=5e-15)
    s = ph.State(m["N"], 0)
    steps = 2000
    m["src"] = ph.source_table(dict(spec, steps=steps))
    Hprev = s.H.copy()
    energies = []
    for t in range(steps):
        ph.step(m, s, t)
        if t > 900:     # source has decayed to < 1e-30
            energies.append(np.sum(s.E ** 2) + np.sum(s.H * Hprev))
        Hprev = s.H.copy()
    e = np.array(energies)
    assert e[0] > 1e-3
    assert np.max(np.abs(e - e[0])) <= 1e-12 * e[0]


def _spectra(spec, freqs):
    f, R, T, _, p0, p1 = ph.spectra(spec, freqs)
    X0 = ph.dft(p0, spec["Sc"] * spec["dz"] / C0, freqs)
    band = np.abs(X0[0]) >= 0.05 * np.abs(X0[0]).max()
    return R, T, band


def test_lossless_dielectric_slab_energy_and_fabry_perot():
    eps = 2.25
    spec = ph.default_spec(N=2400, steps=16000, slab_start=1000, slab_stop=1300, eps_inf=eps, ade=False,
                           src_tau=8e-15, src_t0=40e-15, probes=[60, 2400 - 42])
    freqs = np.linspace(150e12, 250e12, 201)
    R, T, band = _spectra(spec, freqs)
    assert band.sum() > 40
    assert np.max(np.abs(R[band] ** 2 + T[band] ** 2 - 1)) < 1e-5
    # Fabry-Perot: R vanishes exactly where k_s L = m pi, with the medium's numerical
    # dispersion sin(w dt/2) = (Sc/n) sin(k_s dz/2) and L = 300 cells
    dt = spec["Sc"] * spec["dz"] / C0
    n_s = math.sqrt(eps)
    fm = [(2 / dt) * math.asin(spec["Sc"] / n_s * math.sin(m * math.pi / (2 * 300))) / (2 * math.pi)
          for m in range(1, 60)]
    fm = [f for f in fm if 175e12 < f < 225e12]
    assert len(fm) >= 4
    Rm, Tm, _ = _spectra(spec, np.array(fm))
    assert np.max(Rm) < 2e-4
    # half-way between resonances R is at its Fabry-Perot maximum 2r/(1+r^2), r=(n-1)/(n+1)
    r = (n_s - 1) / (n_s + 1)
    fh = [0.5 * (a + b) for a, b in zip(fm[:-1], fm[1:])]
    Rh, _, _ = _spectra(spec, np.array(fh))
    assert np.all(np.abs(Rh - 2 * r / (1 + r * r)) < 2e-2)


def test_lorentz_ade_second_order_convergence():
    w0, g, de = 2 * math.pi * 150e12, 2 * math.pi * 5e12, 2.0
    w = 2 * math.pi * 200e12
    chi = de * w0 ** 2 / (w0 ** 2 - w ** 2 - 1j * g * w)
    errs = []
    base_dt = 0.5 * 1.5e-6 / 40 / C0
    for div in (1, 4):
        dt = base_dt / div
        A, B, Cc = ph.ade_coefficients(dict(gamma=g), dt, np.array([w0]), np.array([de]))
        steps = int(3e-12 / dt)
        P = Pm = 0.0
        tt = np.arange(steps) * dt
        for n in range(steps):   # P^{n+1} from E^n = exp(-i w t_n), complex drive
            Pn = A[0] * P + B[0] * Pm + Cc[0] * np.exp(-1j * w * tt[n])
            Pm, P = P, Pn
        chi_num = P / np.exp(-1j * w * (tt[-1] + dt))     # P^{n+1} against E at t_{n+1}
        errs.append(abs(chi_num - chi) / abs(chi))
    assert errs[0] < 5e-2
    ratio = errs[0] / errs[1]
    assert 12 < ratio < 20     # ~16 for second order


def test_lorentz_slab_absorbs():
    spec = ph.default_spec(N=2400, steps=14000, slab_start=1000, slab_stop=1200, probes=[60, 2400 - 42])
    freqs = np.linspace(120e12, 280e12, 161)
    R, T, band = _spectra(spec, freqs)
    loss = 1 - R[band] ** 2 - T[band] ** 2
    assert np.all(loss > -1e-6) and loss.max() > 1e-3


def test_cpml_reflection_small():
    N = 1200
    spec, m = _vacuum_model(N, 3000, probes=[600], src_cell=300)
    _, pr = ph.run(m)
    specL, mL = _vacuum_model(N + 4000, 3000, probes=[600], src_cell=300)
    _, prL = ph.run(mL)
    refl = np.max(np.abs(pr[0] - prL[0]))
    assert np.max(np.abs(prL[0])) > 0.3
    assert refl < 1e-5 * np.max(np.abs(prL[0]))


def test_kerr_state_satisfies_constitutive_law():
    spec = ph.default_spec(N=900, steps=1600, slab_start=400, slab_stop=500, chi3=2e-2, src_amp=2.0,
                           src_cell=100, src_t0=60e-15, src_tau=15e-15)
    m = ph.build(spec)
    s, _ = ph.run(m)
    E = s.E[400:500]
    res = m["chi3"] * E ** 3 + m["eps"] * E - (s.D - s.P)
    assert np.max(np.abs(E)) > 0.1
    assert np.max(np.abs(res)) <= 1e-13 * np.max(np.abs(s.D - s.P))
    # and the nonlinearity matters at this amplitude
    assert np.max(m["chi3"] * E ** 2 / m["eps"]) > 1e-3


def test_c_restatement_bitwise_equal():
    from oracle import cpu_ref
    cpu_ref.build_lib()
    rng = np.random.default_rng(7)
    ns = 300
    spec = ph.default_spec(N=1500, steps=2500, slab_start=700, slab_stop=700 + ns, chi3=1e-2, src_amp=2.0,
                           eps_inf=rng.uniform(1, 3, ns), omega0=2 * math.pi * 150e12 * (1 + rng.uniform(-.05, .05, ns)))
    m = ph.build(spec)
    s, pr = ph.run(m)
    sim, _ = cpu_ref.run_model(m, threads=1)
    for a, b in ((s.E, sim.E), (s.H, sim.H), (s.D, sim.D), (s.P, sim.P), (s.psiE, sim.psiE),
                 (s.psiH, sim.psiH), (pr, sim.probe_out)):
        assert np.array_equal(a, b)
