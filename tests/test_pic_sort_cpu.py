"""CPU check of the index arithmetic of the PIC kernels (pic_kernels.hpp:
pic_push_count and pic_scatter), emulated lane by lane over several steps of
random moves: the new indices form the stable sort by new cell with parked
particles keeping their index, start[] stays the exclusive count scan, J'
equals the CIC sum of every live particle's two weights (np.bincount), and
the single-writer live range equals the min/max of the live cells.  For the
one-cell-group path, which writes the new state from push/count and stages
the movers into another wave's cells, two ping-pong buffers are followed
through the steps."""
import numpy as np
import pytest

from emulator import emulate_multi_cell_writes, emulate_one_cell_writes, emulate_pic_step, mc_groups


def _moves(rng, cell, N):
    live = (cell >= 2) & (cell <= N - 3)
    d = rng.integers(-1, 2, len(cell))
    newc = np.where(live, cell + d, cell)
    newc = np.where(live & (newc < 2), 1, newc)
    newc = np.where(live & (newc > N - 3), N - 2, newc)
    return np.where(live, newc - cell, 0), newc


@pytest.mark.parametrize("N,n,lo,hi,G,steps", [
    (200, 3000, 2, 198, 1, 4),       # dense, two chunks per group
    (200, 3000, 2, 198, 4, 3),
    (300, 500, 100, 140, 7, 4),      # clustered, empty regions
    (64, 2000, 2, 62, 64, 3),        # one group covers everything
    (500, 40, 2, 498, 16, 4),        # sparse: chunks span many cells
    (120, 900, 2, 6, 3, 6),          # drifting into the left graveyard
    (60, 3000, 54, 58, 1, 6),        # most of the beam parked in the right graveyard
])
def test_pic_step_indices(N, n, lo, hi, G, steps):
    rng = np.random.default_rng(N * 7 + n + G)
    cell = np.sort(rng.integers(lo, hi, n))
    ids = np.arange(n)
    start = np.searchsorted(cell, np.arange(N + 2), side="left").astype(np.int64)
    live = (cell >= 2) & (cell <= N - 3)
    crange = (int(cell[live].min()), int(cell[live].max()))
    urange = crange
    for step in range(steps):
        live = (cell >= 2) & (cell <= N - 3)
        d, newc = _moves(rng, cell, N)
        off = rng.integers(0, 3, n)
        a0, a1 = rng.normal(size=n), rng.normal(size=n)
        new_idx, nstart, J, nrange, nurange, cup, cdown = emulate_pic_step(
            cell, d, off, a0, a1, start, crange, urange, G, N)
        assert np.array_equal(np.sort(new_idx), np.arange(n)), step
        assert np.all(new_idx[~live] == np.nonzero(~live)[0])
        order = np.argsort(new_idx)
        assert np.array_equal(order, np.argsort(newc, kind="stable")), step
        newcell = newc[order]
        assert np.array_equal(nstart, np.searchsorted(newcell, np.arange(N + 2), side="left")), step
        node0 = cell + off - 1
        Jref = np.bincount(node0[live], weights=a0[live], minlength=N + 3)[:N + 1] + \
            np.bincount(node0[live] + 1, weights=a1[live], minlength=N + 3)[:N + 1]
        np.testing.assert_allclose(J, Jref, rtol=0, atol=1e-12)
        nl = (newcell >= 2) & (newcell <= N - 3)
        exp = (int(newcell[nl].min()), int(newcell[nl].max())) if nl.any() else (N, -1)
        assert nrange == exp, (step, nrange, exp)
        cell, ids, start, crange, urange = newcell, ids[order], nstart, nrange, nurange


def _one_cell_pairs(start, cmin, cmax, N, np_):
    """The particle ranges push/count's one-cell path reads, two adjacent
    cells per wave (pic_kernels.hpp, G == 1): each pair loads three cell
    starts (lanes 0-2) at clamped cells and reads them back with readlane."""
    clampi = lambda v: min(max(v, 0), np_)
    cl = lambda c: min(max(c, 2), N - 2)
    lo_cell, hi_cell = max(cmin - 1, 0), min(cmax + 1, N)
    out = []
    for gi in range((hi_cell - lo_cell + 2) >> 1):
        gA = lo_cell + 2 * gi
        gB = gA + 1
        hasB = gB <= hi_cell
        sb = [start[cl(gA)], start[cl(gB)], start[cl(gB + 1 if hasB else gB)]]
        pA = clampi(sb[0])
        pB = max(pA, clampi(sb[1]))
        pE = max(pB, clampi(sb[2]))
        out.append((gA, pA, pB))
        if hasB:
            out.append((gB, pB, pE))
        else:
            assert pE == pB
    return out


@pytest.mark.parametrize("N,lo,hi,n", [(64, 2, 62, 900), (64, 2, 5, 300), (64, 58, 62, 300), (200, 90, 91, 50),
                                       (200, 40, 160, 7)])
def test_one_cell_pairs_cover_each_cell_once(N, lo, hi, n):
    """Pairs of one-cell groups visit every cell of [cmin-1, cmax+1] once, in
    order, each with the particle range the one-cell-per-wave path reads
    (start[max(c, 2)] .. start[min(c+1, N-2)]); graveyard cells read nothing."""
    rng = np.random.default_rng(N + lo + hi + n)
    for park in (0, 1):
        cell = np.sort(np.r_[rng.integers(lo, hi, n), np.full(5 * park, 1), np.full(7 * park, N - 2)])
        start = np.searchsorted(cell, np.arange(N + 2), side="left").astype(np.int64)
        live = (cell >= 2) & (cell <= N - 3)
        cmin, cmax = int(cell[live].min()), int(cell[live].max())
        got = _one_cell_pairs(start, cmin, cmax, N, len(cell))
        lo_cell, hi_cell = max(cmin - 1, 0), min(cmax + 1, N)
        assert [g for g, _, _ in got] == list(range(lo_cell, hi_cell + 1))
        for g, p0, p1 in got:
            plo = start[max(g, 2)]
            phi = max(plo, start[min(g + 1, N - 2)])
            assert (p0, p1) == (plo, phi), g
            assert np.all(cell[p0:p1] == g)
        # every live particle is read exactly once
        assert sum(p1 - p0 for _, p0, p1 in got) == int(live.sum())


@pytest.mark.parametrize("G", [1, 2, 5, 64])
@pytest.mark.parametrize("N,n,lo,hi,steps", [
    (200, 3000, 2, 198, 5),          # dense: up to ~40 per cell
    (80, 9000, 20, 40, 4),           # cells of several hundred (the two-pass case)
    (300, 500, 100, 140, 5),         # clustered, empty regions
    (120, 900, 2, 6, 8),             # drifting into the left graveyard
    (60, 3000, 54, 58, 8),           # most of the beam parked in the right graveyard
    (64, 1, 30, 31, 3),              # one particle
])
def test_write_path_buffers(N, n, lo, hi, steps, G):
    """Two ping-pong buffers through several steps of random moves on the
    one-cell path (G = 1) and the multi-cell path: after every step the
    buffer written holds every particle at its sorted index (the indices of
    emulate_pic_step), parked particles included, although the long-parked
    ones are never rewritten."""
    rng = np.random.default_rng(N * 11 + n)
    cell = np.sort(rng.integers(lo, hi, n))
    start = np.searchsorted(cell, np.arange(N + 2), side="left").astype(np.int64)
    prev_start = start.copy()
    live = (cell >= 2) & (cell <= N - 3)
    crange = (int(cell[live].min()), int(cell[live].max()))
    urange = crange
    buf = [np.arange(n), np.full(n, -1)]          # particle ids; the other buffer holds garbage
    ids = np.arange(n)
    for step in range(steps):
        par = step & 1
        d, newc = _moves(rng, cell, N)
        off = np.ones(n, dtype=np.int64)
        zero = np.zeros(n)
        new_idx, nstart, _, nrange, nurange, cup_ref, cdown_ref = emulate_pic_step(
            cell, d, off, zero, zero, start, crange, urange, max(G, 1), N)
        if G == 1:
            dst, cup, cdown = emulate_one_cell_writes(cell, d, start, prev_start, crange, N)
        else:
            dst, cup, cdown = emulate_multi_cell_writes(cell, d, start, prev_start, crange, N, G)
        lo_cell, hi_cell = max(crange[0] - 1, 0), min(crange[1] + 1, N)
        if crange[0] <= crange[1]:
            sl = slice(lo_cell, hi_cell + 1)
            assert np.array_equal(cup[sl], cup_ref[sl]) and np.array_equal(cdown[sl], cdown_ref[sl])
        w = dst >= 0
        out = buf[par ^ 1].copy()
        out[w] = buf[par][dst[w]]
        # every live particle and every particle parked this or the previous
        # step is written, at its new index
        order = np.argsort(new_idx)
        assert np.array_equal(out, buf[par][order]), step
        assert np.array_equal(new_idx[dst[w]], np.nonzero(w)[0])
        buf[par ^ 1] = out
        prev_start, start = start, nstart
        cell, crange, urange = newc[order], nrange, nurange
        ids = ids[order]
        assert np.array_equal(buf[par ^ 1], ids)


@pytest.mark.parametrize("lo,hi,G", [(1, 3, 2), (1, 4, 2), (5, 70, 64), (5, 69, 64), (2, 200, 7), (10, 12, 64)])
def test_multi_cell_groups_tile_the_range(lo, hi, G):
    """Groups cover [lo, hi] once, in order, each with at least two cells
    (a one-cell remainder joins the last group), at most G + 1."""
    g = mc_groups(lo, hi, G)
    assert g[0][0] == lo and g[-1][1] == hi + 1
    assert all(a[1] == b[0] for a, b in zip(g, g[1:]))
    assert all(2 <= b - a <= G + 1 for a, b in g)
