"""CPU emulation of the device stencil's per-cell semantics (test infrastructure).

Runs an fdtd_pic_amd.engine.SimSpec with the arithmetic the HIP kernel
applies to each cell (include/fdtdpic.h), whole-grid and FP64, so the
descriptor builders can be checked against the oracles without a GPU.
It is never used by the product and never stands in for the device in a
`gpu` test.
"""
import numpy as np

SLAB_NONE, SLAB_DIEL, SLAB_LINEAR, SLAB_KERR_REF, SLAB_KERR_E = range(5)


def _cubic_root0(a, b, c, d):
    from oracle.cubic import root0_1728
    return root0_1728(a, b, c, d)


def _kerr_e(u, e, chi3):
    from oracle.cubic import kerr_root_e
    return kerr_root_e(u, e, chi3)


def run(spec, J=None):
    s = spec
    N, T = s.N, s.steps
    n = N + 1
    full = lambda scal, strip: _expand(scal, strip, s.Ls, s.Rs, n)
    ke = full(s.ke0, s.s_ke)
    kh = full(s.kh0, s.s_kh)
    bE, cE, CbE = (full(0.0, a) for a in (s.s_bE, s.s_cE, s.s_CbE))
    bH, cH, C2H = (full(0.0, a) for a in (s.s_bH, s.s_cH, s.s_C2H))
    strip = np.zeros(n, dtype=bool)
    strip[:s.Ls] = True
    strip[s.Rs:] = True
    F, R = s.F, s.R
    if s.slab_kind == SLAB_DIEL:
        sl = np.zeros(n, dtype=bool)
        sl[F:R] = True
        ke = np.where(sl & ~strip, s.keS, ke)
        kh = np.where(sl & ~strip, s.khS, kh)
    dform = s.slab_kind >= SLAB_LINEAR
    E, H, psE, psH = (np.zeros(n) for _ in range(4))
    ns = R - F if dform else 0
    D, P, Pm = (np.zeros(ns) for _ in range(3))
    probes = np.zeros((len(s.probes), T))
    snaps = np.zeros((s.snap_rows, n)) if s.snap_rows else None
    A = np.asarray(s.sl_A) if s.ade else None
    B = np.asarray(s.sl_B) if s.ade else None
    Cc = np.asarray(s.sl_C) if s.ade else None
    e = np.asarray(s.sl_e) if dform else None
    for t in range(T):
        Jv = np.zeros(n) if J is None else J[t]
        dH = H - np.r_[0.0, H[:-1]]
        if dform:
            if s.ade:
                Pn = A * P + B * Pm + Cc * E[F:R]
                Pm, P = P, Pn
            D = D + (dH[F:R] - s.jD * Jv[F:R]) * s.kD
        Enew = E + (dH - Jv) * ke
        psE = np.where(strip, bE * psE + cE * dH, psE)
        Enew = np.where(strip, Enew - CbE * psE, Enew)
        if dform:
            if s.slab_kind == SLAB_LINEAR:
                Enew[F:R] = (D - P) / e
            elif s.slab_kind == SLAB_KERR_E:
                Enew[F:R] = [_kerr_e(u, ei, s.kerr_chi3) for u, ei in zip(D - P, e)]
            else:
                Acub = np.array([_cubic_root0(s.kerr_a, s.kerr_b, s.kerr_c, -(abs(d / s.kerr_e0)) ** 2)
                                 for d in D])
                Enew[F:R] = D / (s.kerr_e0 * s.kerr_lin + s.kerr_e0 * s.kerr_chi3 * Acub)
        Enew[s.src_cell] += s.srcE[t]
        E = Enew
        if s.tfsf:
            H[s.src_cell - 1] -= s.srcH[t]
        dE = np.r_[E[1:], 0.0] - E
        H = H + dE * kh
        psH = np.where(strip, bH * psH + cH * dE, psH)
        H = np.where(strip, H + C2H * psH, H)
        for k, (c, t0, t1) in enumerate(s.probes):
            if t0 <= t <= t1:
                probes[k, t] = E[c]
        if snaps is not None and s.snap_interval and t > 0 and t % s.snap_interval == 0:
            r = t // s.snap_interval
            if r < s.snap_rows:
                snaps[r] = E
    out = dict(E=E, H=H, PSIE=np.where(strip, psE, 0.0), PSIH=np.where(strip, psH, 0.0), probes=probes)
    if dform:
        out.update(D=D, P=P, PM=Pm)
    if snaps is not None:
        out["snap"] = snaps
    return out


def _expand(scalar, strip_vals, Ls, Rs, n):
    a = np.full(n, float(scalar))
    if strip_vals is not None:
        sv = np.asarray(strip_vals, dtype=np.float64)
        a[:Ls] = sv[:Ls]
        a[Rs:] = sv[Ls:]
    return a
