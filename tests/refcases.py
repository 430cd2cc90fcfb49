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


def fresnel_spec():
    """A Lorentz half-space for the reflection check: config-1 slab material
    from cell 1000 to 100 cells short of the right end of a 9,000-cell grid,
    16,000 steps (the back face's echo would need >= 33,000 steps to return),
    and a 5 fs Gaussian pulse that covers 105-300 THz."""
    from fdtd_pic_amd import configs
    return configs.base(9000, 16000, slab_start=1000, slab_stop=8900, src_tau=5e-15, src_t0=30e-15)


def discrete_fresnel(p, freqs):
    """|r| of a vacuum / Lorentz interface on the Yee grid, in closed form.

    The ADE recurrence P^{n+1} = A P^n + B P^{n-1} + C E^n with
    E^{n+1} = (D^{n+1} - P^{n+1})/eps_inf has, for E ~ z^n, z = exp(i w dt),
    eps_num = eps_inf + C z / (1 - A z - B z^2).  Eliminating H, the grid obeys
    E[j+1] - 2 E[j] + E[j-1] = -(2 sin(w dt/2)/Sc)^2 eps_j E[j], so each side
    carries waves with sin(k/2) = sqrt(eps) sin(w dt/2)/Sc, and matching the
    two solutions on the nodes either side of the interface gives
    r = (exp(-i k1) - exp(-i k0)) / (exp(i k0) - exp(-i k1)); it tends to
    Fresnel's (1 - n)/(1 + n) as the cells shrink.  Also returns the
    continuum Fresnel |r| with the Lorentz eps(w) (what the reference plots,
    BaseFDTD11.py:882-923)."""
    import math
    import numpy as np
    w = 2 * math.pi * np.asarray(freqs, dtype=np.float64)
    dt = p["Sc"] * p["dz"] / C0
    g = p["gamma"]
    w0 = p["omega0"]
    Dd = 1.0 / dt ** 2 + g / (2.0 * dt)
    A = (2.0 / dt ** 2 - w0 ** 2) / Dd
    B = (g / (2.0 * dt) - 1.0 / dt ** 2) / Dd
    Cc = p["d_eps"] * w0 ** 2 / Dd
    z = np.exp(1j * w * dt)
    eps = p["eps_inf"] + Cc * z / (1 - A * z - B * z * z)
    s = np.sin(w * dt / 2) / p["Sc"]
    k0 = 2 * np.arcsin(s)
    k1 = 2 * np.arcsin(np.sqrt(eps.astype(complex)) * s)
    k1 = np.where(k1.imag < 0, -k1, k1)
    r = (np.exp(-1j * k1) - np.exp(-1j * k0)) / (np.exp(1j * k0) - np.exp(-1j * k1))
    n = np.sqrt(p["eps_inf"] + p["d_eps"] * w0 ** 2 / (w0 ** 2 - w ** 2 - 1j * g * w))
    return np.abs(r), np.abs((1 - n) / (1 + n))


def looped_start(kind, freq=3e9, domainSize=0.15, lo=400, hi=600):
    """(V, P, C_V, C_P) before MasterController.LoopedSim, built as
    MasterController.__Main__ builds them (:620-681) for a Lorentz run with a
    sine source and TFSF; kind "oracle" or "product"."""
    from oracle import faithful as FO
    import MasterController as MC
    if kind == "oracle":
        P = FO.make_params(FO.env_setup(freq, domainSize, lo, hi), False, domainSize, freq, 20, LorentzMed=True,
                           SineCont=True, TFSF=True)
        V = FO.make_variables(P.Nz, P.timeSteps, P.vidInterval, 1)
        return V, P, FO.make_cpml_variables(P.Nz, P.timeSteps), FO.make_cpml_params(P.dz)
    import Environment_Setup
    P = MC.Params(*Environment_Setup.envSetup(freq, domainSize, lo, hi), False, domainSize, freq, 20,
                  LorentzMed=True, SineCont=True, TFSF=True)
    V = MC.Variables(P.Nz, P.timeSteps, P.vidInterval, 1)
    return V, P, MC.CPML_Variables(P.Nz, P.timeSteps), MC.CPML_Params(P.dz)


def oracle_looped_sim(points, domainSize=0.15, lo=400, hi=600, Interval=1e8):
    """The reference's "Input frequency sweep" (MasterController.py:543-559)
    with the oracle's functions, one point after another: returns the
    measured and analytical reflection per point and each point's objects."""
    import numpy as np
    from oracle import faithful as FO
    V, P, C_V, C_P = looped_start("oracle", domainSize=domainSize, lo=lo, hi=hi)
    ref, anal, objs = np.zeros(points), np.zeros(points), []
    for k in range(points):
        setup = FO.env_setup(P.freq_in, domainSize, lo, hi, VExists=True, V=V, P=P)
        P = FO.make_params(setup, P.MORmode, domainSize, P.freq_in, 20, LorentzMed=P.LorentzMed,
                           SineCont=P.SineCont, Gaussian=P.Gaussian, TFSF=P.TFSF)
        V = FO.make_variables(P.Nz, P.timeSteps, P.vidInterval, 1)
        C_P = FO.make_cpml_params(P.dz)
        C_V = FO.make_cpml_variables(P.Nz, P.timeSteps)
        FO.controller(V, P, C_V, C_P)
        ref[k] = FO.ref_tester(P, V.x1ColAf) / FO.ref_tester(P, V.x1ColBe)
        anal[k] = FO.analytical_reflection(V, P)
        objs.append((V, P, C_V, C_P))
        P.freq_in = P.freq_in + Interval
    return ref, anal, objs
