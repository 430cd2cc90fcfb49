"""BaseFDTD11's per-step functions called one at a time (the drop-in runs
each as a device kernel) against the oracle's restatement of the same
functions (oracle/faithful.py), on random reference-shaped state: bitwise
equal, except the Kerr root.  The reference's 1728.org Cardano form cancels
(S + U - b/3a), so both the device and the host results sit ~1e-11 from the
exact root (checked to 30 digits: 8.5e-12 and 2.7e-12 on the worst cell) and
their difference is the rounding of pow; that op is held to the FP64 bar,
1e-10 relative."""
import numpy as np
import pytest

from refcases import small_setup, oracle_objects, product_objects

pytestmark = pytest.mark.gpu

V_ARRAYS = ("Ex", "Hy", "Jx", "Dx", "polarisationCurr", "tempTempVarPol", "Acubic")
C_ARRAYS = ("psi_Ex", "psi_Hy", "beX", "ceX", "Cb", "bmY", "cmY", "C2", "den_Exdz", "den_Hydz")


def _pair(seed, **flags):
    from oracle import faithful as FO
    import BaseFDTD11
    setup = small_setup(Nz=300, T=10, W=20, F=150, R=260)
    Vo, Po, Co, _ = oracle_objects(setup, 9e9, **flags)
    Vp, Pp, Cp, CPp = product_objects(setup, 9e9, **flags)
    FO.field_init(Vo, Po)
    FO.empty_space_calc(Vo, Po)
    BaseFDTD11.FieldInit(Vp, Pp)
    BaseFDTD11.EmptySpaceCalc(Vp, Pp)
    rng = np.random.default_rng(seed)
    n = Po.Nz + 1
    for name in V_ARRAYS:
        a = rng.normal(size=n) * (1e-9 if name in ("Dx", "polarisationCurr", "tempTempVarPol") else 1.0)
        setattr(Vo, name, a.copy())
        setattr(Vp, name, a.copy())
    for name in ("UpExMat", "UpHyMat", "UpHySelf"):
        a = rng.uniform(0.5, 2.0, n)
        setattr(Vo, name, a.copy())
        setattr(Vp, name, a.copy())
    for name in C_ARRAYS:
        a = rng.normal(size=n)
        setattr(Co, name, a.copy())
        setattr(Cp, name, a.copy())
    Vo.chi3Stat = Vp.chi3Stat = 0.3
    return (Vo, Po, Co), (Vp, Pp, Cp, CPp)


CASES = [
    ("ADE_ExUpdate", "ade_ex_update", ("Ex",), ()),
    ("ADE_HyUpdate", "ade_hy_update", ("Hy",), ()),
    ("CPML_Psi_e_Update", "cpml_psi_e", ("Ex",), ("psi_Ex",)),
    ("CPML_Psi_m_Update", "cpml_psi_m", ("Hy",), ("psi_Hy",)),
    ("ADE_DxUpdate", "ade_dx_update", ("Dx",), ()),
    ("ADE_ExCreate", "ade_ex_create", ("Ex",), ()),
    ("ADE_PolarisationCurrent_Ex", "ade_polarisation", ("polarisationCurr",), ()),
    ("NonLinExUpdate", "nonlin_ex_update", ("Ex",), ()),
]


@pytest.mark.parametrize("name,oname,vout,cout", CASES)
def test_single_op_bitwise(name, oname, vout, cout):
    from oracle import faithful as FO
    import BaseFDTD11
    (Vo, Po, Co), (Vp, Pp, Cp, CPp) = _pair(hash(name) % 1000, LorentzMed=True)
    fo = getattr(FO, oname)
    args = {"ade_ex_create": (Vo, Po), "ade_polarisation": (Vo, Po), "nonlin_ex_update": (Vo, Po)}.get(
        oname, (Vo, Po, Co))
    fo(*args)
    fn = getattr(BaseFDTD11, name)
    if name in ("NonLinExUpdate",):
        ret = fn(Vp, Pp)
    elif name == "ADE_PolarisationCurrent_Ex":
        ret = fn(Vp, Pp, Cp, CPp, 0)
    elif name == "ADE_ExUpdate":
        ret = fn(Vp, Pp, Cp, CPp, 0)
    else:
        ret = fn(Vp, Pp, Cp, CPp)
    for a in vout:
        assert np.array_equal(getattr(Vp, a), getattr(Vo, a)), (name, a)
    for a in cout:
        assert np.array_equal(getattr(Cp, a), getattr(Co, a)), (name, a)
    if cout:
        assert ret[0] is getattr(Cp, cout[0]) and ret[1] is getattr(Vp, vout[0])
    else:
        assert ret is getattr(Vp, vout[0])


def test_acubic_finder():
    from oracle import faithful as FO
    import BaseFDTD11
    (Vo, Po, Co), (Vp, Pp, Cp, CPp) = _pair(11, nonLinMed=True)
    Vo.Dx[:] = Vp.Dx[:] = np.random.default_rng(3).normal(size=Po.Nz + 1) * 1e-9
    FO.acubic_finder(Vo, Po)
    BaseFDTD11.AcubicFinder(Vp, Pp)
    F, R = Po.materialFrontEdge, Po.materialRearEdge
    ref = Vo.Acubic[F:R]
    got = Vp.Acubic[F:R]
    rel = np.abs(got - ref) / np.maximum(np.abs(ref), 1e-300)
    k = int(np.argmax(rel))
    print("acubic max rel", rel[k], "at", k, "got", repr(got[k]), "ref", repr(ref[k]), "Dx", repr(Vo.Dx[F + k]),
          "coefs", FO.cubic_coefficients(Vo, Po), flush=True)
    assert np.max(np.abs(ref)) > 0
    np.testing.assert_allclose(got, ref, rtol=1e-10, atol=1e-300)
