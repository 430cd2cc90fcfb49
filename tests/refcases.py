"""Small reference-shaped parameter sets shared by the faithful-mode tests.

envSetup's own grids are large (PML = 6*Nlam cells), so the tests build the
20-tuple Params expects directly, with the same derived quantities.
"""
C0 = 299792458.0


def small_setup(Nz=400, T=600, W=20, F=250, R=399, nzsrc=60, freq=9e9, Nlam=40, x1Loc=None, x2Loc=None):
    lamMin = C0 / freq
    dz = lamMin / Nlam
    delT = (dz / C0) * 0.95
    courantNo = (C0 * delT) / dz
    period = 1 / freq
    x1Loc = F - 20 if x1Loc is None else x1Loc
    x2Loc = nzsrc - 25 if x2Loc is None else x2Loc
    return (Nz, T, 0, 0, 1.0, 1.0, 1.0, 1.0, x1Loc, x2Loc, F, R, W, nzsrc, lamMin, dz, delT, courantNo,
            period, Nlam)


def oracle_objects(setup, freq, **flags):
    from oracle import faithful as FO
    P = FO.make_params(setup, False, 0.7, freq, 20, **flags)
    V = FO.make_variables(P.Nz, P.timeSteps, P.vidInterval, 3)
    C_P = FO.make_cpml_params(P.dz)
    C_V = FO.make_cpml_variables(P.Nz, P.timeSteps)
    return V, P, C_V, C_P


def product_objects(setup, freq, **flags):
    import MasterController as MC
    P = MC.Params(*setup, False, 0.7, freq, 20, **flags)
    V = MC.Variables(P.Nz, P.timeSteps, P.vidInterval, 3)
    C_P = MC.CPML_Params(P.dz)
    C_V = MC.CPML_Variables(P.Nz, P.timeSteps)
    return V, P, C_V, C_P


def rel_l2(a, b):
    import numpy as np
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    nb = np.linalg.norm(b)
    if nb == 0:
        return float(np.linalg.norm(a))
    return float(np.linalg.norm(a - b) / nb)
