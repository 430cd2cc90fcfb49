"""CPU check of the sort/deposit index arithmetic of pic_sort_deposit
(pic.hip), emulated lane by lane: for random cell-sorted beams with moves of
-1/0/+1 cell and absorption into the graveyard cells, the new indices form
the stable sort by new cell, the new start[] is the exclusive count scan, and
J' equals the CIC sum of every particle's two weights (np.bincount)."""
import numpy as np
import pytest

from emulator import emulate_pic_sort_deposit


def _beam(rng, N, n, dense_lo, dense_hi, p_dead=0.05):
    cell = np.sort(rng.integers(dense_lo, dense_hi, n))
    dead = rng.random(n) < p_dead
    cell = np.where(dead, np.where(rng.random(n) < 0.5, 1, N - 2), cell)
    cell = np.sort(cell)
    live = (cell >= 2) & (cell <= N - 3)
    d = np.where(live, rng.integers(-1, 2, n), 0)
    newc = cell + d
    newc = np.where(live & (newc < 2), 1, newc)
    newc = np.where(live & (newc > N - 3), N - 2, newc)
    d = newc - cell
    off = np.where(live, rng.integers(0, 3, n), 1)
    code = (d + 1) | (off << 2)
    a0 = np.where(live, rng.normal(size=n), 0.0)
    a1 = np.where(live, rng.normal(size=n), 0.0)
    start = np.searchsorted(cell, np.arange(N + 2), side="left").astype(np.int64)
    return cell, newc, code, a0, a1, off, start


@pytest.mark.parametrize("N,n,lo,hi,G", [
    (200, 3000, 2, 198, 4),        # dense, many chunks per wave
    (200, 3000, 2, 198, 1),
    (300, 500, 100, 140, 7),       # clustered, empty regions
    (64, 2000, 2, 62, 61),         # one wave covers everything
    (500, 40, 2, 498, 16),         # sparse: chunks span many cells
    (120, 900, 2, 5, 3),           # packed against the left graveyard
])
def test_sort_deposit_indices(N, n, lo, hi, G):
    rng = np.random.default_rng(N * 7 + n + G)
    cell, newc, code, a0, a1, off, start = _beam(rng, N, n, lo, hi)
    new_idx, nstart, J = emulate_pic_sort_deposit(cell, code, a0, a1, start, G, N)
    assert np.array_equal(np.sort(new_idx), np.arange(n))
    order = np.argsort(new_idx)
    assert np.array_equal(order, np.argsort(newc, kind="stable"))
    assert np.array_equal(nstart[:N + 1], np.searchsorted(np.sort(newc), np.arange(N + 1), side="left"))
    node0 = cell + off - 1
    Jref = np.bincount(node0, weights=a0, minlength=N + 2)[:N + 1] + \
        np.bincount(node0 + 1, weights=a1, minlength=N + 3)[:N + 1]
    np.testing.assert_allclose(J, Jref, rtol=0, atol=1e-12)
