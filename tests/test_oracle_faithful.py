"""Pins for oracle/faithful.py (and the drop-in parameter builders).

* envSetup on the reference's own call (MasterController.py:636) gives the
  values computed independently: Nlam=350, PML 2100, Nz 11,555, slab 8,404
  cells, and 19,250 steps: freq_in*7000*delT evaluates to exactly 19.0 in
  IEEE double, so the bin search stops at N = 7000 (a hand calculation that
  assumes no exact bin gets the fallback maxim-1 = 7999 and 21,997 instead);
  the drop-in Environment_Setup returns the identical tuple.
* jitclass float32 members round as numba stores them.
* CPML arrays carry the reference's structure: Cb/C2 hole at N+1-W, H
  profiles alias E (bmY == beX), cmY == ceX/dz, den == 1 for kappaMax = 1.
* The faithful Yee update, Lorentz slab and Kerr slab (chi3 -> 0) equal the
  physical model (itself pinned by closed forms in test_oracle_physical.py)
  in normalised units H' = Z0 H, D' = D/eps0 while no wave has reached a PML;
  the only difference allowed is the CharImp-vs-eps0 quirk (~1e-8).
* spatialStab's correction tends to 1 with resolution and is applied
  cumulatively per pass.
"""
import math

import numpy as np
import pytest

from oracle import faithful as FO
from oracle import physical as ph
from refcases import small_setup, oracle_objects, rel_l2


def test_env_setup_reference_call():
    s = FO.env_setup(9e9, 0.7, 7000, 8000, nonLinMed=True)
    Nz, T, F, R, W, Nlam = s[0], s[1], s[10], s[11], s[12], s[19]
    assert (Nlam, W, Nz, R - F, T) == (350, 2100, 11555, 8404, 7000 + int(7000 * 350 / 200))
    dz = (299792458.0 / 9e9) / 350
    assert (9e9 * 7000) / (1 / ((dz / 299792458.0) * 0.95)) == 19.0
    import Environment_Setup
    assert Environment_Setup.envSetup(9e9, 0.7, 7000, 8000, nonLinMed=True) == s


def test_env_setup_variants_match_dropin():
    import Environment_Setup
    import MasterController as MC
    for args, kw in [((10e9, 0.5, 400, 600), {}), ((3e9, 1.0, 400, 600), {}), ((9e9, 0.7, 500, 501), {})]:
        assert Environment_Setup.envSetup(*args, **kw) == FO.env_setup(*args, **kw)
    P = MC.Params(*FO.env_setup(10e9, 0.5), False, 0.5, 10e9, 20)
    V = MC.Variables(P.Nz, P.timeSteps, 50, 1)
    Po = FO.make_params(FO.env_setup(10e9, 0.5), False, 0.5, 10e9, 20)
    Vo = FO.make_variables(Po.Nz, Po.timeSteps, 50, 1)
    assert Environment_Setup.envSetup(10e9, 0.5, VExists=True, V=V, P=P) == \
        FO.env_setup(10e9, 0.5, VExists=True, V=Vo, P=Po)


def test_mat_setup_matches_dropin():
    import Material_Def
    assert Material_Def.matSetup(None, None, None, None, 5e9, 600) == FO.mat_setup(None, None, None, None, 5e9, 600)


def test_float32_members():
    import MasterController as MC
    setup = FO.env_setup(9e9, 0.7, 7000, 8000, nonLinMed=True)
    P = MC.Params(*setup, False, 0.7, 9e9, 20)
    assert P.courantNo == float(np.float32(setup[17])) and P.courantNo != setup[17]
    assert MC.Variables(10, 100, 50, 1).alpha3 == 0.699999988079071
    assert MC.CPML_Params(1e-4).alphaMax == float(np.float32(0.05))
    assert P.domainSize == 0 and isinstance(P.Nz, int)
    Po = FO.make_params(setup, False, 0.7, 9e9, 20)
    for k in ("courantNo", "delT", "dz", "Nz", "timeSteps", "materialFrontEdge", "pmlWidth", "period"):
        assert getattr(P, k) == getattr(Po, k), k


def test_cpml_structure():
    setup = small_setup(Nz=300, T=10, W=20, F=200, R=299)
    V, P, C_V, C_P = oracle_objects(setup, 9e9, LorentzMed=True)
    FO.field_init(V, P)
    FO.empty_space_calc(V, P)
    FO.bound_cond_manager(V, P, C_V, C_P)
    n, W = P.Nz + 1, P.pmlWidth
    assert C_V.Cb[n - W] == 0 and C_V.C2[n - W] == 0          # the hole
    assert np.all(C_V.Cb[:W] == V.UpExMat[:W]) and np.all(C_V.Cb[n - W + 1:] == V.UpExMat[n - W + 1:])
    assert np.array_equal(C_V.bmY, C_V.beX)
    np.testing.assert_allclose(C_V.cmY[:W], C_V.ceX[:W] / P.dz, rtol=1e-15)
    assert np.all(C_V.den_Exdz == 1) and np.all(C_V.den_Hydz == 1)
    assert np.all(C_V.beX[:W] < 1) and np.all(C_V.ceX[:W] < 0)
    # the drop-in builders produce the same arrays
    import MasterController as MC
    import BaseFDTD11
    import Solver_Engine
    Pp = MC.Params(*setup, False, 0.7, 9e9, 20, LorentzMed=True)
    Vp = MC.Variables(Pp.Nz, Pp.timeSteps, 50, 1)
    Cp = MC.CPML_Variables(Pp.Nz, Pp.timeSteps)
    BaseFDTD11.FieldInit(Vp, Pp)
    BaseFDTD11.EmptySpaceCalc(Vp, Pp)
    BaseFDTD11.CPML_FieldInit(Vp, Pp, Cp, MC.CPML_Params(Pp.dz))
    Solver_Engine.boundCondManager(Vp, Pp, Cp, MC.CPML_Params(Pp.dz))
    for k in ("beX", "ceX", "bmY", "cmY", "Cb", "C2", "den_Exdz", "den_Hydz"):
        np.testing.assert_array_equal(getattr(Cp, k), getattr(C_V, k), err_msg=k)


def _faithful_loop(V, P, C_V, steps, src, slab=None, nl=False):
    """Faithful per-step kernels, single pass, no TFSF (oracle functions)."""
    for t in range(steps):
        if slab == "lorentz":
            FO.ade_temp_pol_curr(V)
            FO.ade_polarisation(V, P)
        FO.ade_ex_update(V, P, C_V)
        FO.cpml_psi_e(V, P, C_V)
        V.Ex[P.nzsrc] += src[t]
        if slab:
            FO.ade_dx_update(V, P, C_V)
            if nl:
                FO.acubic_finder(V, P)
                FO.nonlin_ex_update(V, P)
            else:
                FO.ade_ex_create(V, P)
        FO.ade_hy_update(V, P, C_V)
        FO.cpml_psi_m(V, P, C_V)


def _physical_twin(P, steps, src, F, R, eps=1.0, ade=None, chi3=0.0):
    spec = ph.default_spec(N=P.Nz, steps=steps, Sc=P.courantNo, dz=P.dz, slab_start=F, slab_stop=R,
                           eps_inf=eps, chi3=chi3, ade=ade is not None, src_cell=P.nzsrc, probes=[])
    m = ph.build(spec)
    m["src"] = np.asarray(src)
    if ade is not None:
        A, B, Cc = ade
        ns = R - F
        m["A"], m["B"], m["C"] = np.full(ns, A), np.full(ns, B), np.full(ns, Cc)
    s = ph.State(m["N"], R - F)
    for t in range(steps):
        ph.step(m, s, t)
    return s


def _setup_objects(Nz, slab, **flags):
    setup = small_setup(Nz=Nz, T=1, W=20, F=slab[0], R=slab[1], nzsrc=Nz // 4)
    V, P, C_V, C_P = oracle_objects(setup, 9e9, **flags)
    FO.field_init(V, P)
    FO.empty_space_calc(V, P)
    FO.bound_cond_manager(V, P, C_V, C_P)
    return V, P, C_V


def _src(P, steps):
    t = np.arange(steps)
    return np.exp(-((t - 60) / 15.0) ** 2) * np.sin(2 * np.pi * t / 40.0)


def test_faithful_vacuum_equals_physical_yee():
    V, P, C_V = _setup_objects(2000, (1990, 1995))
    steps = 400
    src = _src(P, steps)
    _faithful_loop(V, P, C_V, steps, src)
    s = _physical_twin(P, steps, src, 1990, 1990)
    assert np.max(np.abs(V.Ex)) > 0.1
    assert rel_l2(V.Ex, s.E) < 1e-12
    assert rel_l2(V.Hy * P.CharImp, s.H) < 1e-12


def test_faithful_lorentz_slab_equals_physical_ade():
    V, P, C_V = _setup_objects(2000, (700, 1200), LorentzMed=True)
    V.plasmaFreqE, V.omega_0E, V.gammaE = 1.5 * 2 * np.pi * 9e9, 2 * np.pi * 9.5e9, 2 * np.pi * 0.5e9
    steps = 450         # the left-going wave stays clear of the (sign-flipped) left CPML
    src = _src(P, steps)
    _faithful_loop(V, P, C_V, steps, src, slab="lorentz")
    A, B, Cc = FO.ade_coefficients(V, P)
    s = _physical_twin(P, steps, src, 700, 1200, eps=1.0, ade=(A, B, Cc / P.permit_0))
    assert np.max(np.abs(V.polarisationCurr)) > 0
    assert rel_l2(V.Ex, s.E) < 1e-6             # dt/(eps0 dz) vs CharImp*Sc quirk ~1e-8/step
    assert rel_l2(V.polarisationCurr / P.permit_0, np.r_[np.zeros(700), s.P, np.zeros(801)]) < 1e-6


def test_faithful_nl_slab_linear_limit_equals_dielectric():
    V, P, C_V = _setup_objects(1600, (700, 1000), nonLinMed=True)
    V.chi3Stat = 0.0
    steps = 380
    src = _src(P, steps)
    _faithful_loop(V, P, C_V, steps, src, slab="kerr", nl=True)
    s = _physical_twin(P, steps, src, 700, 1000, eps=math.sqrt(1.2))
    assert rel_l2(V.Ex, s.E) < 1e-6


def test_spatial_stab_limit_and_cumulative():
    pf, w0, g = np.sqrt(1.5) * 2 * np.pi * 20e9, 2 * np.pi * 20e9, 2 * np.pi * 2e9
    fixes = []
    for Nlam in (40, 400, 4000):
        dz = (299792458.0 / 9e9) / Nlam
        dt = dz / 299792458.0 * 0.95
        fixes.append(FO.spatial_stab(100, 10 * Nlam, dz, 9e9, dt, pf, w0, g)[4])
    errs = [abs(abs(f) - 1) for f in fixes]
    assert errs[2] < errs[1] < errs[0] and errs[2] < 1e-4
    # the integrator applies it once per pass (Solver_Engine.py:286): twice for LinLor
    setup = small_setup(Nz=300, T=50, F=200, R=299)
    V, P, C_V, C_P = oracle_objects(setup, 9e9, LorentzMed=True, SineCont=True)
    pf0 = V.plasmaFreqE
    FO.controller(V, P, C_V, C_P)
    f1 = FO.spatial_stab(P.timeSteps, P.Nz, P.dz, P.freq_in, P.delT, pf0, V.omega_0E, V.gammaE)[3]
    f2 = FO.spatial_stab(P.timeSteps, P.Nz, P.dz, P.freq_in, P.delT, f1, V.omega_0E, V.gammaE)[3]
    assert V.plasmaFreqE == f2 != f1


def test_guards_exit():
    import BaseFDTD11
    import MasterController as MC
    setup = small_setup(Nz=300, T=10, F=200, R=299)
    P = MC.Params(*setup, False, 0.7, 9e9, 20)
    V = MC.Variables(P.Nz, P.timeSteps, 50, 1)
    P.Nz = 30000
    with pytest.raises(SystemExit):
        BaseFDTD11.FieldInit(V, P)
    with pytest.raises(SystemExit):
        FO.field_init(V, P)
