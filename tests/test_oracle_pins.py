"""Closed-form pins for the oracle functions that were checked only against a
restatement by the same author (the product's twin) or not at all.

Every test states the mutation it catches, and the `test_planted_*` tests
plant that mistake in the oracle (on the CPU, by monkeypatching) and show
the pin rejects it.  The expected values never come from the function under
test: they are the physics or arithmetic the function must reproduce.

PIC coupling (oracle/pic.py + the J' term of oracle/physical.step; the
reference's sink is V.Jx in ADE_ExUpdate, BaseFDTD11.py:663-669):
  * a smoothly ramped current sheet J' radiates E = -J'/2 on both sides and
    H' = +-J'/2 (the plane-wave jump condition [H'] = J', E continuous);
  * a cold beam's deposit sums to Z0 n q v_x L (its surface current times
    Z0), computed from density, extent and drift, not from jscale;
  * the CIC deposit of a uniformly moving particle is centred at t^{n+1/2};
  * the gather averages H^{n-1/2} and H^{n+1/2}.
Faithful set-up (oracle/faithful.py and the drop-ins BaseFDTD11.py /
Solver_Engine.py):
  * CPML psi recursions with the difference frozen follow the geometric sum
    psi_n = c dH (1 - b^n)/(1 - b), at the Cb/C2 hole as well;
  * the CPML arrays invert to the polynomial grading sigma_opt ((W-k)/W)^4,
    alpha_max ((k+1)/W) (BaseFDTD11.py:222-296, MasterController.py:349-360);
  * SmoothTurnOn/SourceManager tables are continuous-time sines at the
    float32 Courant number, with the 0.8 f pump at 0.1 (E) and 0.01 (H);
  * the cubic coefficients' root x satisfies |eps(w) + a3 chi3 x|^2 x =
    |D/eps0|^2 with the Lorentz eps(w) evaluated independently, and the NL
    E update solves E (eps0 sqrt(1.2) + eps0 chi3 A) = D.
"""
import math

import numpy as np
import pytest

from oracle import faithful as FO
from oracle import physical as ph
from oracle import pic as P
from refcases import oracle_objects, product_objects, small_setup

Z0 = 376.730313668


# ---------------------------------------------------------------------------
# PIC coupling units
# ---------------------------------------------------------------------------
def _sheet_fields(step, N=2000, n0=1000, J0=1.0, ramp=400, steps=1600, probe=400):
    """Vacuum grid with a prescribed current sheet J'(t) at node n0, ramped
    by a raised cosine over `ramp` steps; returns E and H' at n0 +- probe."""
    spec = ph.default_spec(N=N, steps=steps, slab_start=n0, slab_stop=n0, src_amp=0.0, probes=[])
    m = ph.build(spec)
    s = ph.State(N, 0)
    J = np.zeros(N + 1)
    for t in range(steps):
        J[n0] = J0 * (0.5 - 0.5 * math.cos(math.pi * min(t, ramp) / ramp))
        step(m, s, t, J)
    return s.E[n0 - probe], s.E[n0 + probe], s.H[n0 - probe], s.H[n0 + probe - 1]


def _sheet_ok(f, J0=1.0, tol=1e-5):
    El, Er, Hl, Hr = f
    return (abs(El + J0 / 2) < tol * J0 and abs(Er + J0 / 2) < tol * J0 and abs(Hl + J0 / 2) < tol * J0
            and abs(Hr - J0 / 2) < tol * J0)


def test_current_sheet_radiates_minus_half_J():
    """Catches: J' entering E with the wrong sign, without the Courant factor,
    or through the H update (E = -J'/2, H' = -/+J'/2 left/right)."""
    assert _sheet_ok(_sheet_fields(ph.step))


def test_planted_current_sign_fails_sheet_pin():
    def wrong(m, s, t, J=None):          # J' with the sign flipped
        ph.step(m, s, t, None if J is None else -J)
    assert not _sheet_ok(_sheet_fields(wrong))

    def no_courant(m, s, t, J=None):     # J' added without the Sc factor
        ph.step(m, s, t, None if J is None else J / m["Sc"])
    assert not _sheet_ok(_sheet_fields(no_courant))


def _cold_beam(vx, n_part=4000, density=1e16, lo=300, hi=340, N=600, dz=37.5e-9, Sc=0.5):
    """Uniform cold sheet with a transverse drift vx (no motion along z)."""
    rng = np.random.default_rng(11)
    x = lo + rng.uniform(0, hi - lo, n_part)
    beta = np.zeros((n_part, 3))
    beta[:, 0] = vx / P.C0
    u = beta / np.sqrt(1 - np.sum(beta ** 2, axis=1))[:, None]
    p = dict(beam=dict(cell_lo=lo, cell_hi=hi, density=density))
    q, m, w = P.beam_params(p, n_part, dz)
    dt = Sc * dz / P.C0
    return P.Beam(x, u, q, m, w, dt, N, Sc), (hi - lo) * dz


def test_beam_surface_current_from_density_and_drift():
    """Catches: jscale without w or c, u_x used for beta_x (gamma = 1.048 at
    0.3c), or a weight from the wrong extent: sum J' = Z0 n q v_x L."""
    vx, n = 0.3 * P.C0, 1e16
    b, L = _cold_beam(vx, density=n)
    z = np.zeros(b.N + 1)
    J = P.push_deposit(b, z, z, z)
    expect = Z0 * n * (-P.E_CHARGE) * vx * L
    assert abs(J.sum() - expect) < 1e-12 * abs(expect)
    # the product's beam inputs give the same weight (configs.beam_inputs)
    from fdtd_pic_amd import configs
    pp = configs.base(600, 10)
    pp["beam"] = dict(seed=1, n=4000, cell_lo=300, cell_hi=340, v_drift=0.0, v_th=1e-3 * P.C0, density=n)
    _, _, q, m, w = configs.beam_inputs(pp)
    assert abs(Z0 * q * w * P.C0 * 4000 - Z0 * n * q * P.C0 * L) < 1e-12 * abs(Z0 * n * q * P.C0 * L)


def test_planted_u_for_beta_fails_surface_current_pin():
    """The surface-current pin rejects a deposit of q w c u_x (u_x for beta_x)."""
    vx, n = 0.3 * P.C0, 1e16
    b, L = _cold_beam(vx, density=n)
    b.jscale *= 1.0 / math.sqrt(1.0 - 0.09)        # q w c beta_x gamma = q w c u_x
    z = np.zeros(b.N + 1)
    J = P.push_deposit(b, z, z, z)
    expect = Z0 * n * (-P.E_CHARGE) * vx * L
    assert abs(J.sum() - expect) > 1e-3 * abs(expect)


def _node_series(push, k=200, nstar=40, steps=80):
    """J' at node k over the steps of one particle crossing it at dx = 0.25
    cells per step, aligned so that x^{n*+1/2} = k."""
    Sc = 0.5
    # beta_z = 0.5 exactly: dx = Sc beta_z = 0.25; u_x = 0.1 carries the current
    ux = 0.1
    uz = math.sqrt(0.25 * (1 + ux * ux) / 0.75)
    b = P.Beam([k - (nstar + 0.5) * 0.25], [[ux, 0.0, uz]], q=-P.E_CHARGE, m=P.M_E, w=1.0,
               dt=Sc * 37.5e-9 / P.C0, N=400, Sc=Sc)
    z = np.zeros(401)
    out = []
    for _ in range(steps):
        out.append(push(b, z, z, z)[k])
    return np.array(out)


def test_deposit_centred_at_half_step():
    """Catches: depositing at x^n (centroid n* + 1/2) or x^{n+1} (n* - 1/2):
    the node's current series is a triangle centred on the step whose
    half-step position is the node."""
    J = _node_series(P.push_deposit)
    n = np.arange(len(J))
    assert abs((n * J).sum() / J.sum() - 40) < 1e-9
    assert np.count_nonzero(np.abs(J) > 1e-6 * np.abs(J).max()) == 7    # |x - k| < 1: 7 quarter-cell steps


def test_planted_deposit_at_old_position_fails():
    def at_old(b, E, Hp, Hc):            # deposit at x^n: shift the half-step position back by dx/2
        b.x -= 0.125
        J = P.push_deposit(b, E, Hp, Hc)
        b.x += 0.125
        return J
    J = _node_series(at_old)
    n = np.arange(len(J))
    assert abs((n * J).sum() / J.sum() - 40) > 0.4


def test_gather_averages_the_two_H_levels():
    """Catches: gathering only H^{n+1/2} or only H^{n-1/2} (the two differ here)."""
    N = 100
    k = np.arange(N + 1, dtype=np.float64)
    Hp = 3.0 - 0.5 * (k + 0.5)            # H' node k sits at x = k + 1/2
    Hc = -1.0 + 0.25 * (k + 0.5)
    x = np.array([10.0, 10.25, 10.5, 33.9, 50.49])
    _, Hg = P.gather(x, np.zeros(N + 1), Hp, Hc)
    np.testing.assert_allclose(Hg, 0.5 * ((3.0 - 0.5 * x) + (-1.0 + 0.25 * x)), rtol=1e-14)


# ---------------------------------------------------------------------------
# faithful CPML
# ---------------------------------------------------------------------------
def _cpml_sets(W=20, Nz=300):
    setup = small_setup(Nz=Nz, T=10, W=W, F=200, R=299)
    V, Pp, C_V, C_P = oracle_objects(setup, 9e9, LorentzMed=True)
    FO.field_init(V, Pp)
    FO.empty_space_calc(V, Pp)
    FO.bound_cond_manager(V, Pp, C_V, C_P)
    import BaseFDTD11
    import Solver_Engine
    Vd, Pd, Cd, CPd = product_objects(setup, 9e9, LorentzMed=True)
    BaseFDTD11.FieldInit(Vd, Pd)
    BaseFDTD11.EmptySpaceCalc(Vd, Pd)
    BaseFDTD11.CPML_FieldInit(Vd, Pd, Cd, CPd)
    Solver_Engine.boundCondManager(Vd, Pd, Cd, CPd)
    return (V, Pp, C_V), (Vd, Pd, Cd)


def _psi_e_closed_form(C_V, dH, n):
    b, c = C_V.beX, C_V.ceX
    with np.errstate(invalid="ignore", divide="ignore"):
        geo = np.where(b != 1, (1 - b ** n) / (1 - b), n)
        # sum_{k=1..n} (1 - b^k)/(1 - b) = (n - b (1 - b^n)/(1 - b)) / (1 - b)
        cum = np.where(b != 1, (n - b * geo) / (1 - b), n * (n + 1) / 2)
    return c * dH * geo, c * dH * cum


def test_psi_e_recursion_geometric_sum_at_hole():
    """Catches: beX/ceX swapped, Ex += Cb psi (the reference subtracts,
    BaseFDTD11.py:371,375), or the hole at n-W moved (psi updated there, Ex not)."""
    for (V, Pp, C_V) in _cpml_sets():
        n, W = Pp.Nz + 1, Pp.pmlWidth
        rng = np.random.default_rng(5)
        V.Hy = rng.normal(0, 1, n)
        dH = np.r_[0.0, V.Hy[1:] - V.Hy[:-1]]
        Ex0 = rng.normal(0, 1, n)
        V.Ex = Ex0.copy()
        C_V.psi_Ex = np.zeros(n)
        steps = 7
        for _ in range(steps):
            FO.cpml_psi_e(V, Pp, C_V)
        psi, cum = _psi_e_closed_form(C_V, dH, steps)
        cells = np.r_[1:W, n - W:n]
        np.testing.assert_allclose(C_V.psi_Ex[cells], psi[cells], rtol=1e-12, atol=1e-300)
        np.testing.assert_allclose(V.Ex[cells], (Ex0 - C_V.Cb * cum)[cells], rtol=1e-12)
        hole = n - W
        assert C_V.psi_Ex[hole] != 0 and V.Ex[hole] == Ex0[hole] and C_V.Cb[hole] == 0
        rest = np.setdiff1d(np.arange(n), cells)
        assert np.all(C_V.psi_Ex[rest] == 0) and np.array_equal(V.Ex[rest], Ex0[rest])


def test_psi_m_recursion_geometric_sum():
    """Catches: H -= C2 psi, cmY without /dz, or psi_H run over n-1 (the
    right strip ends at n-2, BaseFDTD11.py:388)."""
    for (V, Pp, C_V) in _cpml_sets():
        n, W = Pp.Nz + 1, Pp.pmlWidth
        rng = np.random.default_rng(6)
        V.Ex = rng.normal(0, 1, n)
        dE = np.r_[V.Ex[1:] - V.Ex[:-1], 0.0]
        H0 = rng.normal(0, 1, n)
        V.Hy = H0.copy()
        C_V.psi_Hy = np.zeros(n)
        steps = 5
        for _ in range(steps):
            FO.cpml_psi_m(V, Pp, C_V)
        b, c = C_V.bmY, C_V.cmY
        geo = (1 - b ** steps) / (1 - b)
        cum = (steps - b * geo) / (1 - b)
        cells = np.r_[1:W, n - W:n - 1]
        np.testing.assert_allclose(C_V.psi_Hy[cells], (c * dE * geo)[cells], rtol=1e-12)
        np.testing.assert_allclose(V.Hy[cells], (H0 + C_V.C2 * c * dE * cum)[cells], rtol=1e-12)
        assert C_V.psi_Hy[n - 1] == 0 and V.Hy[n - 1] == H0[n - 1]
        assert np.allclose(c[:W - 1] * Pp.dz, C_V.ceX[:W - 1], rtol=1e-14)


def test_cpml_arrays_invert_to_polynomial_grading():
    """Catches: a wrong grading power, k off by one, an unmirrored right
    strip, or sigma_opt off its MasterController.py:356 value: with kappa = 1,
    -ln(b) eps0/dt = sigma + alpha and c/(b - 1) = sigma/(sigma + alpha)."""
    eps0 = 8.8541878188e-12
    mu0 = 1.25663706127e-06
    for (V, Pp, C_V) in _cpml_sets():
        n, W = Pp.Nz + 1, Pp.pmlWidth
        strips = np.r_[0:W, n - W:n]
        tot = np.zeros(n)
        sig = np.zeros(n)
        tot[strips] = -np.log(C_V.beX[strips]) * eps0 / Pp.delT
        sig[strips] = C_V.ceX[strips] / (C_V.beX[strips] - 1) * tot[strips]
        alp = tot - sig
        k = np.arange(W)
        sig_opt = 0.5 * 0.8 / (Pp.dz * math.sqrt(mu0 / eps0))
        a_max = float(np.float32(0.05))
        np.testing.assert_allclose(sig[:W], sig_opt * ((W - k) / W) ** 4, rtol=1e-9)
        np.testing.assert_allclose(alp[:W], a_max * (k + 1) / W, rtol=1e-6)
        np.testing.assert_allclose(sig[n - W:], (sig_opt * ((W - k) / W) ** 4)[::-1], rtol=1e-9)
        np.testing.assert_allclose(alp[n - W:], (a_max * (k + 1) / W)[::-1], rtol=1e-6)
        assert np.all(C_V.beX[W:n - W] == 0)


# ---------------------------------------------------------------------------
# faithful sources
# ---------------------------------------------------------------------------
def _continuous_tables(P_, nonlin):
    sc = float(np.float32(P_.courantNo))          # the float32 jitclass member
    n = np.arange(P_.timeSteps, dtype=np.float64)
    on = n * P_.delT < P_.period * P_.Periods
    tE, tH = n * sc * P_.dz / 299792458.0, (n + 1) * sc * P_.dz / 299792458.0
    f = P_.freq_in
    E = np.where(on, np.sin(2 * np.pi * f * tE), 0.0) * sc
    H = np.where(on, np.sin(2 * np.pi * f * tH), 0.0) * sc
    if nonlin:
        E = E + 0.1 * sc * np.where(on, np.sin(2 * np.pi * 0.8 * f * tE), 0.0)
        H = H + 0.01 * sc * np.where(on, np.sin(2 * np.pi * 0.8 * f * tH), 0.0)
    if P_.TFSF:
        H = H / Z0
    return E, H


@pytest.mark.parametrize("nonlin", [False, True])
def test_source_tables_are_continuous_time_sines(nonlin):
    """Catches: the phase at float64 Sc (1e-8 relative, ~1e-6 absolute here),
    the pump amplitudes swapped (0.1 E / 0.01 H), Hys without its one-step
    lead, or the TFSF H table not divided by Z0."""
    import Solver_Engine
    flags = dict(SineCont=True, TFSF=True, nonLinMed=nonlin, LorentzMed=not nonlin)
    setup = small_setup(Nz=300, T=900, F=200, R=299)
    _, Po, _, _ = oracle_objects(setup, 9e9, **flags)
    Vd, Pd, Cd, CPd = product_objects(setup, 9e9, **flags)
    for P_ in (Po, Pd):
        P_.Periods = 10                            # the table switches off inside the run
    assert Pd.courantNo != setup[17]               # float32 rounding is in play
    E, H = _continuous_tables(Po, nonlin)
    assert np.count_nonzero(E) < len(E) - 100
    Eo, Ho = FO.source_manager(Po)
    Ed, Hd = Solver_Engine.SourceManager(Vd, Pd, Cd, CPd)
    for a, b in ((Eo, E), (Ed, E)):
        np.testing.assert_allclose(np.asarray(a), b, rtol=0, atol=1e-12)
    for a, b in ((Ho, H), (Hd, H)):
        np.testing.assert_allclose(np.asarray(a), b, rtol=0, atol=1e-12 / Z0)


def test_planted_float64_courant_fails_source_pin():
    setup = small_setup(Nz=300, T=900, F=200, R=299)
    _, Po, _, _ = oracle_objects(setup, 9e9, SineCont=True, LorentzMed=True)
    Po.Periods = 10
    E, _ = _continuous_tables(Po, False)
    exact = float(setup[17])
    n = np.arange(Po.timeSteps)
    on = n * Po.delT < Po.period * Po.Periods
    wrong = np.where(on, np.sin(2 * np.pi * Po.freq_in * n * exact * Po.dz / 299792458.0), 0.0) * exact
    assert np.max(np.abs(wrong - E)) > 1e-9


# ---------------------------------------------------------------------------
# faithful Kerr coefficients and update
# ---------------------------------------------------------------------------
def _lorentz_eps(wp, w0, g, f):
    w = 2 * math.pi * f
    # 1 + wp^2 / ((w0 - w)(w0 + w) + i g w), written through its real and
    # imaginary parts
    a, b = (w0 - w) * (w0 + w), g * w
    den = a * a + b * b
    return complex(1 + wp * wp * a / den, -wp * wp * b / den)


def test_cubic_coefficients_root_solves_intensity_equation():
    """Catches: qua without its factor 2, cub without alpha3 squared, or eps
    at w instead of 2 pi f: the root x of AcubicFinder's cubic satisfies
    |eps(w) + a3 chi3 x|^2 x = |D/eps0|^2."""
    import BaseFDTD11
    setup = small_setup(Nz=300, T=10, F=200, R=299)
    Vo, Po, _, _ = oracle_objects(setup, 9e9, nonLinMed=True)
    Vd, Pd, _, _ = product_objects(setup, 9e9, nonLinMed=True)
    for V, P_, coef in ((Vo, Po, FO.cubic_coefficients), (Vd, Pd, BaseFDTD11.cubic_coefficients)):
        cub, qua, one = coef(V, P_)
        eps = _lorentz_eps(V.plasmaFreqE, V.omega_0E, V.gammaE, P_.freq_in)
        a3c = V.alpha3 * V.chi3Stat
        # D/eps0 where the cubic and quadratic terms carry 0.06 % .. 50 % of
        # the sum (the solver's own cancellation limits it to ~1e-7 near D = 0)
        for Dn in (30.0, 100.0, 300.0, 1000.0):
            x = FO.root0_1728(cub, qua, one, -Dn * Dn)
            assert x > 0
            lhs = abs(eps + a3c * x) ** 2 * x
            assert abs(lhs - Dn * Dn) < 1e-9 * Dn * Dn


def test_nl_update_solves_constitutive_relation():
    """Catches: alpha3 in the E update or the Lorentz eps in place of
    sqrt(1.2) (BaseFDTD11.py:858-877 uses neither): E (eps0 sqrt(1.2) +
    eps0 chi3 A) = D with A the intensity root of |D/eps0|."""
    setup = small_setup(Nz=300, T=10, F=200, R=299)
    V, P_, _, _ = oracle_objects(setup, 9e9, nonLinMed=True)
    FO.field_init(V, P_)
    rng = np.random.default_rng(2)
    F, R = P_.materialFrontEdge, P_.materialRearEdge
    V.Dx[F:R] = rng.choice([-1.0, 1.0], R - F) * rng.uniform(30, 1000, R - F) * P_.permit_0
    V.Dx[F:F + 3] = 1e-17                 # |D/eps0|^2 <= 1e-8: A = 0 (BaseFDTD11.py:834)
    FO.acubic_finder(V, P_)
    FO.nonlin_ex_update(V, P_)
    eps0, e = P_.permit_0, math.sqrt(1.2)
    A = V.Acubic[F:R]
    cub, qua, one = FO.cubic_coefficients(V, P_)
    d = np.abs(V.Dx[F:R] / eps0) ** 2
    big = d > 1e-8
    assert big.any() and (~big).any()
    np.testing.assert_allclose((cub * A ** 3 + qua * A ** 2 + one * A)[big], d[big], rtol=1e-9)
    assert np.all(A[~big] == 0)
    np.testing.assert_allclose(V.Ex[F:R] * (eps0 * e + eps0 * V.chi3Stat * A), V.Dx[F:R], rtol=1e-14)


def test_planted_cubic_coefficient_mistakes_fail():
    """The intensity pin rejects qua without its 2 and cub without alpha3^2."""
    setup = small_setup(Nz=300, T=10, F=200, R=299)
    V, P_, _, _ = oracle_objects(setup, 9e9, nonLinMed=True)
    cub, qua, one = FO.cubic_coefficients(V, P_)
    eps = _lorentz_eps(V.plasmaFreqE, V.omega_0E, V.gammaE, P_.freq_in)
    a3c = V.alpha3 * V.chi3Stat
    for c3, c2 in ((cub, qua / 2), (cub / V.alpha3 ** 2, qua)):
        worst = 0.0
        for Dn in (30.0, 100.0, 300.0, 1000.0):
            x = FO.root0_1728(c3, c2, one, -Dn * Dn)
            worst = max(worst, abs(abs(eps + a3c * x) ** 2 * x - Dn * Dn) / (Dn * Dn))
        assert worst > 1e-3


def test_planted_psi_sign_fails_recursion_pin():
    """The psi_E pin rejects Ex += Cb psi (the standard CPML sign, which the
    reference does not use)."""
    (V, Pp, C_V), _ = _cpml_sets()
    n, W = Pp.Nz + 1, Pp.pmlWidth
    rng = np.random.default_rng(5)
    V.Hy = rng.normal(0, 1, n)
    dH = np.r_[0.0, V.Hy[1:] - V.Hy[:-1]]
    Ex0 = rng.normal(0, 1, n)
    V.Ex = Ex0.copy()
    C_V.psi_Ex = np.zeros(n)
    C_V.Cb = -C_V.Cb
    for _ in range(7):
        FO.cpml_psi_e(V, Pp, C_V)
    C_V.Cb = -C_V.Cb
    _, cum = _psi_e_closed_form(C_V, dH, 7)
    cells = np.r_[1:W, n - W:n]
    assert not np.allclose(V.Ex[cells], (Ex0 - C_V.Cb * cum)[cells], rtol=1e-12)
