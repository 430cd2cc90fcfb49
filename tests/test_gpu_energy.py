"""FP64 field-energy accumulator on the device (fdtd_set_energy /
fdtd_read_energy) against the oracle and the leapfrog invariant.

* Closed cavity (no CPML, PEC ends), a pulse injected then switched off:
  W = sum E^2 + sum H' H'prev equals the oracle's W (oracle/physical.py,
  formed the same way) at every recorded step to 1e-12 in FP64 (1e-5 in
  FP32), and once the source has decayed it is constant to 1e-12 of itself
  in FP64 (5e-6 in FP32, whose fields round at ~1e-7 per step).
* With CPML ends the pulse leaves: W drops below 1e-6 of its peak.
* Rows follow the launches in which the simulation steps (every-th one)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _cavity(N=400, steps=2000, Sc=0.5):
    from fdtd_pic_amd import engine
    from oracle import physical as ph
    spec = ph.default_spec(N=N, steps=steps, Sc=Sc, slab_start=N // 2, slab_stop=N // 2, src_cell=150,
                           src_t0=20e-15, src_tau=5e-15, probes=[])
    m = ph.build(spec)
    m["pmlE"][:] = False
    m["pmlH"][:] = False
    s = engine.SimSpec(N, steps)
    ke = np.full(N + 1, Sc)
    ke[0] = ke[N] = 0.0
    kh = np.full(N + 1, Sc)
    kh[N] = 0.0
    z = np.zeros(N + 1)
    s.set_cells(ke, kh, z, z, z, z, z, z, Sc, Sc)
    s.src_cell = 150
    s.srcE = ph.source_table(spec)
    return m, s


def _oracle_energy(m):
    from oracle import physical as ph
    st = ph.State(m["N"], 0)
    W = np.zeros(m["steps"])
    for t in range(m["steps"]):
        Hprev = st.H.copy()
        ph.step(m, st, t)
        W[t] = np.sum(st.E ** 2) + np.sum(st.H * Hprev)
    return W


@pytest.mark.parametrize("prec,tol,cons", [("f64", 1e-12, 1e-12), ("f32", 1e-5, 5e-6)])
def test_cavity_energy_conserved_and_equal_to_oracle(prec, tol, cons):
    from fdtd_pic_amd import engine
    m, spec = _cavity()
    Wo = _oracle_energy(m)
    with engine.Context(prec, K=2, S=16) as ctx:
        sim = ctx.add(spec)
        sim.set_energy(every=1, rows=200)
        ctx.run_all()
        steps, W = sim.energy()
    assert len(steps) == 2000 // 16 and np.array_equal(steps, np.arange(15, 2000, 16))
    np.testing.assert_allclose(W, Wo[steps], rtol=tol)
    late = W[steps > 900]              # the source has decayed below 1e-30
    assert late[0] > 1e-3
    assert np.abs(late - late[0]).max() <= cons * late[0]


def test_energy_every_and_absorbing_ends():
    from fdtd_pic_amd import configs, engine
    p = configs.base(1200, 4000, slab_start=600, slab_stop=600, src_cell=300, probes=[],
                     src_t0=30e-15, src_tau=6e-15)
    with engine.Context("f64", K=2, S=16) as ctx:
        sim = ctx.add(engine.physical_spec(p))
        sim.set_energy(every=5, rows=100)
        ctx.run_all()
        steps, W = sim.energy()
    assert np.array_equal(steps, np.arange(79, 4000, 80))
    assert W.max() > 0 and W[-1] < 1e-6 * W.max()
