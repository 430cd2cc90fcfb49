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
# PIC push/count + scatter (pic.hip: pic_push_count, pic_scatter), lane by lane
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


def emulate_pic_step(cell, d_live, off_live, a0, a1, start, crange, urange, G, N):
    """One step of the two kernels given each live particle's move d and CIC
    record.  Returns (new index per particle, new start, J, next range, next
    union range, cup, cdown)."""
    n = len(cell)
    live = (cell >= 2) & (cell <= N - 3)
    d = np.where(live, d_live, 0)
    off = np.where(live, off_live, 1)
    a0 = np.where(live, a0, 0.0)
    a1 = np.where(live, a1, 0.0)
    cup = np.full(N + 2, -999, dtype=np.int64)      # garbage outside what the kernel writes
    cdown = np.full(N + 2, -999, dtype=np.int64)
    S = np.full((N + 2, 4), np.nan)
    rank = np.full(n, -1, dtype=np.int64)
    cmin, cmax = crange
    nrange = [N, -1]
    nurange = list(urange)
    if cmin <= cmax:
        lo_cell, hi_cell = max(cmin - 1, 0), min(cmax + 1, N)
        ngroups = (hi_cell - lo_cell + G) // G
        for gi in range(ngroups):
            g0 = lo_cell + gi * G
            g1 = min(g0 + G, hi_cell + 1)
            cnt = np.zeros((G, 3), dtype=np.int64)
            Sl = np.zeros((G, 4))
            plo, phi = start[max(g0, 2)], start[min(g1, N - 2)]
            phi = max(plo, phi)
            for base in range(plo, phi, 64):
                i, valid, c, fl, last = _chunk(cell, base, phi)
                ii = np.minimum(i, n - 1)
                dd = np.where(valid, d[ii], 0)
                for l in np.nonzero(valid)[0]:
                    lk = c[l] - g0
                    rank[i[l]] = cnt[lk, dd[l] + 1] + np.sum(dd[fl[l]:l] == dd[l])
                v = np.zeros((64, 4))
                v[np.arange(64), np.where(valid, off[ii], 1)] += np.where(valid, a0[ii], 0.0)
                v[np.arange(64), np.where(valid, off[ii], 1) + 1] += np.where(valid, a1[ii], 0.0)
                for l in np.nonzero(last)[0]:
                    lk = c[l] - g0
                    for q in range(3):
                        cnt[lk, q] += np.sum(dd[fl[l]:l + 1] == q - 1)
                    Sl[lk] += v[fl[l]:l + 1].sum(axis=0)
            for k in range(g1 - g0):
                cup[g0 + k], cdown[g0 + k], S[g0 + k] = cnt[k, 2], cnt[k, 0], Sl[k]
    inr = lambda c: cmin <= cmax and cmin - 1 <= c <= cmax + 1
    up = lambda c: cup[c] if inr(c) else 0
    down = lambda c: cdown[c] if inr(c) else 0
    new_idx = np.empty(n, dtype=np.int64)
    for i in range(n):
        c = cell[i]
        if live[i]:
            r = rank[i]
            if d[i] < 0:
                new_idx[i] = start[c] - up(c - 1) + r
            elif d[i] == 0:
                new_idx[i] = start[c] + down(c) + r
            else:
                new_idx[i] = start[c + 1] - up(c) + down(c + 1) + r
        else:
            new_idx[i] = i
    nstart = start.copy()
    J = np.zeros(N + 1)
    e0, e1 = max(urange[0] - 3, 0), min(urange[1] + 4, N)
    nlo = start[2] + down(2)
    nhi = start[N - 2] - up(N - 3)
    Sq = lambda c, w: S[c, w] if inr(c) else 0.0
    for m in range(e0, e1 + 1):
        J[m] = Sq(m - 2, 3) + Sq(m - 1, 2) + Sq(m, 1) + Sq(m + 1, 0)
        ns = start[m] - up(m - 1) + down(m)
        nstart[m] = ns
        if 2 <= m <= N - 3:
            ns1 = start[m + 1] - up(m) + down(m + 1)
            if ns1 > ns:
                if ns == nlo:
                    nrange[0] = m
                    nurange[0] = min(urange[0], m)
                if ns1 == nhi:
                    nrange[1] = m
                    nurange[1] = max(urange[1], m)
    return new_idx, nstart, J, tuple(nrange), tuple(nurange), cup, cdown


def emulate_one_cell_writes(cell, d_live, start, prev_start, crange, N, waves=4):
    """The writes of one step on the one-cell-group path (pic_kernels.hpp,
    G == 1): push/count's pairs of cells in blocks of `waves` pairs (one
    workgroup), the direct writes and the staging, then the scatter's node
    threads (staged movers, particles parked at the previous step).  Returns
    (dst, cup, cdown): dst[j] = the old index of the particle written to
    index j of the other buffer, -1 where nothing is written.  Asserts that
    every index is in range and written at most once, and that each staged
    slot lies in its cell's old range."""
    n = len(cell)
    live = (cell >= 2) & (cell <= N - 3)
    d = np.where(live, d_live, 0)
    cup = np.zeros(N + 2, dtype=np.int64)
    cdown = np.zeros(N + 2, dtype=np.int64)
    dst = np.full(n, -1, dtype=np.int64)
    staged = np.full(n, -1, dtype=np.int64)

    def put(arr, j, src):
        assert 0 <= j < n, (j, n)
        assert arr[j] == -1, ("written twice", j)
        arr[j] = src

    cmin, cmax = crange
    clampi = lambda v: min(max(int(v), 0), n)
    cl = lambda c: min(max(c, 2), N - 2)
    lo_cell, hi_cell = max(cmin - 1, 0), min(cmax + 1, N)
    if cmin <= cmax:
        npairs = (hi_cell - lo_cell + 2) >> 1
        for bk in range((npairs + waves - 1) // waves):
            pairs = []
            for wv in range(waves):                 # counting, then the exchange
                gi = bk * waves + wv
                pv = gi < npairs
                gA, gB = lo_cell + 2 * gi, lo_cell + 2 * gi + 1
                hasB = pv and gB <= hi_cell
                pA = clampi(start[cl(gA)])
                pB = max(pA, clampi(start[cl(gB)])) if pv else pA
                pE = max(pB, clampi(start[cl(gB + 1 if hasB else gB)])) if pv else pA
                ranks, tot = {}, []
                for p0, p1 in ((pA, pB), (pB, pE)):
                    cnt = [0, 0, 0]
                    for i in range(p0, p1):
                        ranks[i] = cnt[d[i] + 1]
                        cnt[d[i] + 1] += 1
                    tot.append(cnt)
                pairs.append((pv, gA, gB, hasB, pA, pB, pE, ranks, tot))
            for wv, (pv, gA, gB, hasB, pA, pB, pE, ranks, tot) in enumerate(pairs):
                (tA0, _, tA2), (tB0, _, tB2) = tot
                first, last = wv == 0, wv == waves - 1
                up_prev = 0 if first else pairs[wv - 1][8][1][2]
                down_next = 0 if last else pairs[wv + 1][8][0][0]
                for i in range(pA, pB):
                    r = ranks[i]
                    if d[i] < 0 and first:
                        assert pA <= pA + r < pB
                        put(staged, pA + r, i)
                    else:
                        put(dst, pA - up_prev + r if d[i] < 0 else
                            (pA + tA0 + r if d[i] == 0 else pB - tA2 + tB0 + r), i)
                for i in range(pB, pE):
                    r = ranks[i]
                    if d[i] > 0 and last:
                        assert pB <= pB + r < pE
                        put(staged, pB + r, i)
                    else:
                        put(dst, pE - tB2 + down_next + r if d[i] > 0 else
                            (pB + tB0 + r if d[i] == 0 else pB - tA2 + r), i)
                if pv:
                    cup[gA], cdown[gA] = tA2, tA0
                    if hasB:
                        cup[gB], cdown[gB] = tB2, tB0
    inr = lambda c: cmin <= cmax and cmin - 1 <= c <= cmax + 1
    up = lambda c: cup[c] if inr(c) else 0
    down = lambda c: cdown[c] if inr(c) else 0
    moved = set()
    if cmin <= cmax:
        blk = 2 * waves
        for j in range((hi_cell - lo_cell + blk) // blk):
            md = lo_cell + j * blk
            mu = md + blk - 1
            if 2 <= md <= N - 3:
                for r in range(cdown[md]):
                    assert staged[start[md] + r] >= 0
                    put(dst, start[md] - up(md - 1) + r, staged[start[md] + r])
                    moved.add(int(staged[start[md] + r]))
            if 2 <= mu <= min(hi_cell, N - 3):
                for r in range(cup[mu]):
                    assert staged[start[mu] + r] >= 0
                    put(dst, start[mu + 1] - up(mu) + down(mu + 1) + r, staged[start[mu] + r])
                    moved.add(int(staged[start[mu] + r]))
    assert moved == set(staged[staged >= 0].tolist())      # every staged particle moved once
    a0, b1 = prev_start[2], prev_start[N - 2]
    a1, b0 = max(a0, start[2]), min(start[N - 2], b1)
    for i in list(range(a0, a1)) + list(range(b0, b1)):
        put(dst, i, i)
    return dst, cup, cdown


def mc_groups(lo, hi, G):
    """Groups of a multi-cell beam over [lo, hi] (pic_kernels.hpp: mc_groups,
    mc_group): G cells each, the last taking a one-cell remainder."""
    n = hi - lo + 1
    g = (n + G - 1) // G
    ng = g - 1 if (n % G == 1 and g > 1) else g
    return [(lo + j * G, hi + 1 if j == ng - 1 else lo + j * G + G) for j in range(ng)]


def emulate_multi_cell_writes(cell, d_live, start, prev_start, crange, N, G):
    """The writes of one step on the multi-cell path (pic_kernels.hpp,
    pic_push_count and move_block): per group, the totals of a counting
    pass, then each particle's rank and new index in the writing pass; the
    down-movers of a group's first cell and the up-movers of its last are
    staged at start[c] + r and moved by the node blocks.  Returns
    (dst, cup, cdown) as emulate_one_cell_writes does."""
    n = len(cell)
    live = (cell >= 2) & (cell <= N - 3)
    d = np.where(live, d_live, 0)
    cup = np.zeros(N + 2, dtype=np.int64)
    cdown = np.zeros(N + 2, dtype=np.int64)
    dst = np.full(n, -1, dtype=np.int64)
    staged = np.full(n, -1, dtype=np.int64)

    def put(arr, j, src):
        assert 0 <= j < n, (j, n)
        assert arr[j] == -1, ("written twice", j)
        arr[j] = src

    cmin, cmax = crange
    lo_cell, hi_cell = max(cmin - 1, 0), min(cmax + 1, N)
    groups = mc_groups(lo_cell, hi_cell, G) if cmin <= cmax else []
    for g0, g1 in groups:
        m = g1 - g0
        assert m >= 2
        plo = start[max(g0, 2)]
        phi = max(plo, start[min(g1, N - 2)])
        tot = np.zeros((m, 3), dtype=np.int64)
        for i in range(plo, phi):
            tot[cell[i] - g0, d[i] + 1] += 1
        stl = [start[min(g0 + k, N + 1)] for k in range(m + 1)]
        carry = np.zeros((m, 3), dtype=np.int64)
        for i in range(plo, phi):
            k, cls = cell[i] - g0, d[i] + 1
            r = carry[k, cls]
            carry[k, cls] += 1
            if d[i] == 0:
                put(dst, stl[k] + tot[k, 0] + r, i)
            elif d[i] < 0:
                if k == 0:
                    assert stl[k] + r <= i
                    put(staged, stl[k] + r, i)
                else:
                    put(dst, stl[k] - tot[k - 1, 2] + r, i)
            else:
                if k == m - 1:
                    assert stl[k] + r <= i
                    put(staged, stl[k] + r, i)
                else:
                    put(dst, stl[k + 1] - tot[k, 2] + tot[k + 1, 0] + r, i)
        for k in range(m):
            cup[g0 + k], cdown[g0 + k] = tot[k, 2], tot[k, 0]
    inr = lambda c: cmin <= cmax and cmin - 1 <= c <= cmax + 1
    up = lambda c: cup[c] if inr(c) else 0
    down = lambda c: cdown[c] if inr(c) else 0
    moved = set()
    for g0, g1 in groups:
        md, mu = g0, g1 - 1
        if 2 <= md <= N - 3:
            for r in range(cdown[md]):
                put(dst, start[md] - up(md - 1) + r, staged[start[md] + r])
                moved.add(int(staged[start[md] + r]))
        if 2 <= mu <= min(hi_cell, N - 3):
            for r in range(cup[mu]):
                put(dst, start[mu + 1] - up(mu) + down(mu + 1) + r, staged[start[mu] + r])
                moved.add(int(staged[start[mu] + r]))
    assert moved == set(staged[staged >= 0].tolist())
    a0, b1 = prev_start[2], prev_start[N - 2]
    a1, b0 = max(a0, start[2]), min(start[N - 2], b1)
    for i in list(range(a0, a1)) + list(range(b0, b1)):
        put(dst, i, i)
    return dst, cup, cdown
