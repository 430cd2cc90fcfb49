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


# ---------------------------------------------------------------------------
# PIC sort/deposit wave (pic.hip: pic_sort_deposit), lane by lane
# ---------------------------------------------------------------------------
def _chunk(cell, base, hi):
    lanes = np.arange(64)
    i = base + lanes
    valid = i < hi
    c = np.where(valid, cell[np.minimum(i, len(cell) - 1)], 2 ** 31 - 1)
    prev = np.r_[c[0], c[:-1]]
    bnd = valid & ((lanes == 0) | (prev != c))
    fl = np.maximum.accumulate(np.where(bnd, lanes, 0))
    nxt = np.r_[c[1:], -1]
    last = valid & ((lanes == 63) | (i + 1 >= hi) | (nxt != c))
    return i, valid, c, fl, last


def emulate_pic_sort_deposit(cell, code, a0, a1, start, G, N):
    """Returns (new index per old particle, new start[], J) as the kernel computes them."""
    n = len(cell)
    new_idx = np.full(n, -1, dtype=np.int64)
    nstart = start.copy()
    J = np.zeros(N + 1)
    L = G + 3
    for c0 in range(0, N + 1, G):
        cnt = np.zeros((L, 3), dtype=np.int64)
        rk = np.zeros((L, 3), dtype=np.int64)
        S = np.zeros((L, 4))
        cend = min(c0 + G, N + 1)
        kb = c0 - 2
        clamp = lambda v: min(max(int(v), 0), n)
        lo = clamp(start[max(c0 - 2, 0)])
        hi = max(lo, clamp(start[min(c0 + G + 1, N + 1)]))
        for base in range(lo, hi, 64):
            i, valid, c, fl, last = _chunk(cell, base, hi)
            ii = np.minimum(i, n - 1)
            cd = np.where(valid, code[ii], 5)
            d = (cd & 3) - 1
            off = (cd >> 2) & 3
            v = np.zeros((64, 4))
            v[np.arange(64), off] += np.where(valid, a0[ii], 0.0)
            v[np.arange(64), off + 1] += np.where(valid, a1[ii], 0.0)
            for l in np.nonzero(last)[0]:
                seg = slice(fl[l], l + 1)
                lk = c[l] - kb
                if 0 <= lk < L:
                    for dd in range(3):
                        cnt[lk, dd] += np.sum(d[seg] == dd - 1)
                    S[lk] += v[seg].sum(axis=0)
        olo = clamp(start[c0])
        ohi = max(olo, clamp(start[cend]))
        for base in range(olo, ohi, 64):
            i, valid, c, fl, last = _chunk(cell, base, ohi)
            ii = np.minimum(i, n - 1)
            d = np.where(valid, (code[ii] & 3) - 1, 0)
            ranks = np.zeros(64, dtype=np.int64)
            for l in np.nonzero(valid)[0]:
                lk = c[l] - kb
                ranks[l] = rk[lk, d[l] + 1] + np.sum(d[fl[l]:l] == d[l])
            for l in np.nonzero(last)[0]:
                lk = c[l] - kb
                for dd in range(3):
                    rk[lk, dd] += np.sum(d[fl[l]:l + 1] == dd - 1)
            for l in np.nonzero(valid)[0]:
                cc, lk, r = c[l], c[l] - kb, ranks[l]
                if d[l] < 0:
                    ni = start[cc] - cnt[lk - 1, 2] + r
                elif d[l] == 0:
                    ni = start[cc] + cnt[lk, 0] + r
                else:
                    ni = start[cc + 1] - cnt[lk, 2] + cnt[lk + 1, 0] + r
                new_idx[i[l]] = ni
        for cc in range(c0, cend):
            lk = cc - kb
            nstart[cc] = start[cc] - cnt[lk - 1, 2] + cnt[lk, 0]
            if cc <= N:
                J[cc] = S[lk - 2, 3] + S[lk - 1, 2] + S[lk, 1] + S[lk + 1, 0]
    return new_idx, nstart, J
