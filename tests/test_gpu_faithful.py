"""GPU parity of the drop-in integrators against the faithful FP64 oracle.

Each test builds the same reference-shaped parameters twice: once with the
oracle's restatement of the reference (oracle/faithful.py) and once with the
drop-in modules (MasterController.Params/Variables/... ->
Solver_Engine.Integrator* -> libfdtdpic.so), runs MasterController.Controller
on both and compares every array the reference leaves behind.  Tolerance:
relative L2 <= 1e-10 (FP64), as the loop is FP64 in the reference.
"""
import numpy as np
import pytest

from refcases import small_setup, oracle_objects, product_objects, rel_l2

pytestmark = pytest.mark.gpu
TOL = 1e-10
FREQ = 9e9


def _run_both(setup, flags, tweak=None):
    from oracle import faithful as FO
    import MasterController as MC
    Vo, Po, Co, CPo = oracle_objects(setup, FREQ, **flags)
    Vp, Pp, Cp, CPp = product_objects(setup, FREQ, **flags)
    if tweak:
        tweak(Po)
        tweak(Pp)
    print(f"case N={setup[0]} T={setup[1]} flags={flags}", flush=True)
    FO.controller(Vo, Po, Co, CPo)
    MC.Controller(Vp, Pp, Cp, CPp)
    return (Vo, Po, Co), (Vp, Pp, Cp)


def _cmp(o, p, names, tol=TOL):
    errs = {}
    for obj_o, obj_p, name in names:
        errs[name] = rel_l2(getattr(p[obj_p], name), getattr(o[obj_o], name))
    print({k: f"{v:.2e}" for k, v in errs.items()}, flush=True)
    for k, v in errs.items():
        assert v <= tol, f"{k}: rel L2 {v:.3e} > {tol}"


FIELDS = [(0, 0, "Ex"), (0, 0, "Hy"), (2, 2, "psi_Ex"), (2, 2, "psi_Hy"), (0, 0, "x1ColBe"),
          (0, 0, "x1ColAf")]


def test_linlor_two_pass():
    setup = small_setup(Nz=400, T=700)
    o, p = _run_both(setup, dict(LorentzMed=True, SineCont=True, TFSF=True))
    _cmp(o, p, FIELDS + [(0, 0, "Dx"), (0, 0, "polarisationCurr"), (0, 0, "tempVarPol"),
                         (0, 0, "Ex_History")])
    assert np.max(np.abs(p[0].polarisationCurr)) > 0   # the ADE actually ran
    assert o[0].plasmaFreqE == p[0].plasmaFreqE         # cumulative spatialStab rescale


def test_linlor_gaussian_soft_source():
    setup = small_setup(Nz=300, T=500, F=180, R=299, nzsrc=50)

    def tw(P):
        P.Gaussian = True
        P.Amplitude = 0.5
    o, p = _run_both(setup, dict(LorentzMed=True, Gaussian=True, TFSF=False), tw)
    _cmp(o, p, FIELDS + [(0, 0, "Dx"), (0, 0, "polarisationCurr")])


def test_nonlinear_kerr_cubic():
    setup = small_setup(Nz=300, T=400, F=200, R=299, nzsrc=60)

    def tw(P):
        P.FreeSpace = False
        P.Periods = 1000
    o, p = _run_both(setup, dict(nonLinMed=True, SineCont=True, TFSF=True), tw)
    _cmp(o, p, [(0, 0, "Ex"), (0, 0, "Hy"), (2, 2, "psi_Ex"), (2, 2, "psi_Hy"), (0, 0, "Dx"),
                (0, 0, "Port1"), (0, 0, "Port2"), (0, 0, "Ex_History")])
    # the cubic branch was exercised: |Dx/eps0|^2 well above the 1e-8 cut
    assert np.max(np.abs(p[0].Dx / 8.8541878188e-12)) ** 2 > 1e-4


def test_free_space_dielectric_slab():
    setup = small_setup(Nz=360, T=600, F=200, R=300, nzsrc=60)

    def tw(P):
        P.epsRe = 2.25
        P.atten = True
    o, p = _run_both(setup, dict(SineCont=True, TFSF=True), tw)
    _cmp(o, p, FIELDS + [(0, 0, "Ex_History"), (0, 0, "x1Atten")])


def test_free_space_vacuum_no_tfsf():
    setup = small_setup(Nz=300, T=400, F=200, R=299, nzsrc=60)
    o, p = _run_both(setup, dict(SineCont=True, TFSF=False))
    _cmp(o, p, FIELDS)


def test_snapshot_archive_round_trip(tmp_path):
    """Device snapshot rows -> V.Ex_History -> .npz -> back: the oracle's
    Ex_History (vidMake, Solver_Engine.py:57-68) to 1e-10."""
    from fdtd_pic_amd import snapshots
    setup = small_setup(Nz=360, T=600, F=200, R=300, nzsrc=60)
    o, p = _run_both(setup, dict(SineCont=True, TFSF=True))
    f = snapshots.write_history(str(tmp_path / "h.npz"), p[0], p[1])
    d = snapshots.read(f)
    assert np.abs(d["Ex_History"]).max() > 0
    assert rel_l2(d["Ex_History"], o[0].Ex_History) <= TOL
    np.testing.assert_array_equal(d["step"], np.arange(len(o[0].Ex_History)) * p[1].vidInterval)


def test_looped_sim_batched_input_frequency_sweep():
    """MasterController.LoopedSim(loop=True) (MasterController.py:530-569):
    every point planned on the host in the reference's order, then all
    points' passes run as one batch of scheduler jobs on the device.  Each
    point's arrays match the oracle's point-by-point loop to 1e-10; the
    measured reflection results(RefCo) (TransformHandler.RefTester peak
    ratio) to 1e-10 and AnalyticalReflectionE (a host closed form) to
    1e-13."""
    import MasterController as MC
    from refcases import looped_start, oracle_looped_sim
    pts = 4
    ref_o, anal_o, objs_o = oracle_looped_sim(pts)
    V, P, C_V, C_P = looped_start("product")
    freqs, ref, anal, objs = MC.LoopedSim(None, V, P, C_V, C_P, False, 0.15, 400, 600, loop=True, Low=3e9,
                                          Interval=1e8, points=pts)[6]
    print("measured", ref, "analytical", anal, flush=True)
    assert len(objs) == pts and len(freqs) == pts
    np.testing.assert_allclose(ref, ref_o, rtol=1e-10)
    np.testing.assert_allclose(anal, anal_o, rtol=1e-13)     # host closed form (1 ulp across hosts)
    for k in range(pts):
        o = objs_o[k][:3]
        p = objs[k][:3]
        _cmp(o, p, FIELDS + [(0, 0, "Dx"), (0, 0, "polarisationCurr"), (0, 0, "Ex_History")])


def test_main_driver_matches_oracle():
    """MasterController.__Main__ (MasterController.py:620-681): its set-up
    (envSetup at 9 GHz with the Kerr slab (nonLinMed), a sine source with
    TFSF, 1,000 periods, attenAmount 10, vidInterval 50) and LoopedSim's
    single run (loop=False), at a 0.15 m domain and 400-600 steps (5,776
    cells, 1,647 steps; the reference's own values, 0.7 m and 7,000-8,000
    steps, give 11,555 cells and 19,250 steps).  The oracle builds the same
    objects with its restatement and runs its Controller; every array the
    run leaves behind matches to 1e-10."""
    import MasterController as MC
    from oracle import faithful as FO
    V, P, C_P, C_V, Rep = MC.__Main__(freq_in=FREQ, domainSize=0.15, lowLimTim=400, highLimTim=600)
    Po = FO.make_params(FO.env_setup(FREQ, 0.15, 400, 600, nonLinMed=True), False, 0.15, FREQ, 20,
                        nonLinMed=True, SineCont=True, TFSF=True)
    Po.vidInterval = 50
    Vo = FO.make_variables(Po.Nz, Po.timeSteps, Po.vidInterval, 10)
    CPo, Co = FO.make_cpml_params(Po.dz), FO.make_cpml_variables(Po.Nz, Po.timeSteps)
    Po.atten, Po.epsRe, Po.CPMLXp, Po.CPMLXm, Po.TFSF = False, 1, True, True, True
    Po.Gaussian, Po.SineCont, Po.Periods = False, True, 1000
    Po.LorentzMed, Po.nonLinMed, Po.FreeSpace = False, True, False
    print(f"__Main__: N={P.Nz} T={P.timeSteps}", flush=True)
    assert (P.Nz, P.timeSteps) == (Po.Nz, Po.timeSteps) == (5776, 1647)
    FO.controller(Vo, Po, Co, CPo)
    _cmp((Vo, Po, Co), (V, P, C_V), FIELDS + [(0, 0, "Dx"), (0, 0, "Ex_History")])
    assert np.abs(V.Ex).max() > 0.5 and np.abs(V.Dx).max() > 0
