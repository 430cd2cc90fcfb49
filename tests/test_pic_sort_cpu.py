"""CPU check of the index arithmetic of the PIC kernels (pic.hip:
pic_push_count and pic_scatter), emulated lane by lane over several steps of
random moves: the new indices form the stable sort by new cell with parked
particles keeping their index, start[] stays the exclusive count scan, J'
equals the CIC sum of every live particle's two weights (np.bincount), and
the single-writer live range equals the min/max of the live cells."""
import numpy as np
import pytest

from emulator import emulate_pic_step


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
