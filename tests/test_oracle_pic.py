"""Pins for oracle/pic.py from closed forms (the reference has no particle code
to compare with):

* Boris with E only adds q E dt / (m c) to u per step, exactly.
* Boris with B only preserves |u| and turns u by 2 atan(|t|) per step, which
  tends to the relativistic gyro-angle q B dt / (gamma m).
* Crossed uniform E and B (|E| < c|B|) give the E x B drift E/B along z.
* CIC: the grid sum of J' equals jscale * sum(u_x / gamma) of the live
  particles; the deposit is independent of particle order.
* Gather is exact on linear fields (E on integer nodes, H' on half-integer).
* A particle leaving the live cells is parked in cell 1 / N-2 and stops
  depositing.
* A cold beam (u_x = 0) in zero fields deposits nothing and drifts at
  exactly Sc * beta_z cells per step.
"""
import math

import numpy as np

from oracle import pic as P
from oracle import physical as ph


def _beam(x, u, N=400, Sc=0.5, dz=37.5e-9):
    dt = Sc * dz / P.C0
    return P.Beam(x, u, q=-P.E_CHARGE, m=P.M_E, w=3.75e6, dt=dt, N=N, Sc=Sc)


def test_boris_pure_E_exact():
    rng = np.random.default_rng(0)
    u = rng.normal(0, 0.1, (50, 3))
    Eg = rng.normal(0, 1e6, 50)
    a = -1.8e-5
    un = P.boris(u, Eg, np.zeros(50), a)
    np.testing.assert_allclose(un[:, 0], u[:, 0] + 2 * a * Eg, rtol=1e-14, atol=1e-18)
    np.testing.assert_array_equal(un[:, 1:], u[:, 1:])


def test_boris_pure_B_rotation():
    u = np.array([[0.3, 0.1, -0.2]])
    a, H = 2e-6, 3e4                      # t_y = -a H / gamma
    g = math.sqrt(1 + np.sum(u * u))
    ty = -a * H / g
    un = P.boris(u, np.zeros(1), np.full(1, H), a)
    assert abs(np.linalg.norm(un) - np.linalg.norm(u)) < 1e-15
    ang = math.atan2(u[0, 0] * un[0, 2] - u[0, 2] * un[0, 0], u[0, 0] * un[0, 0] + u[0, 2] * un[0, 2])
    assert abs(abs(ang) - 2 * math.atan(abs(ty))) < 1e-13
    # small |t|: the turn tends to q B dt / (gamma m) = 2 a H' / gamma (B = -H'/c,
    # a = q dt / (2 m c)); the difference is O(t^3)
    H = 30.0
    un = P.boris(u, np.zeros(1), np.full(1, H), a)
    ang = math.atan2(u[0, 0] * un[0, 2] - u[0, 2] * un[0, 0], u[0, 0] * un[0, 0] + u[0, 2] * un[0, 2])
    wc = 2 * a * H / g
    assert abs(abs(ang) - wc) < wc * wc * wc


def test_exb_drift():
    a, H, E = 1e-4, 1e3, 10.0            # E / (c B) = E / H' = 0.01
    u = np.zeros((1, 3))
    vs = []
    for _ in range(20000):
        u = P.boris(u, np.full(1, E), np.full(1, H), a)
        g = math.sqrt(1 + np.sum(u * u))
        vs.append(u[0, 2] / g)
    # E x B along z; with B_y = -H'/c the drift beta_z = -E/H'
    assert abs(np.mean(vs) - (-E / H)) < 0.01 * E / H


def test_gather_linear_fields_exact():
    N = 100
    n = np.arange(N + 1, dtype=np.float64)
    E = 2.0 + 0.5 * n
    H = -1.0 + 0.25 * (n + 0.5)           # H' node k sits at x = k + 1/2
    x = np.array([10.0, 10.25, 10.5, 33.9, 50.49])
    Eg, Hg = P.gather(x, E, H, H)
    np.testing.assert_allclose(Eg, 2.0 + 0.5 * x, rtol=1e-15)
    np.testing.assert_allclose(Hg, -1.0 + 0.25 * x, rtol=1e-14)


def test_cic_sum_and_order_independence():
    rng = np.random.default_rng(3)
    N = 400
    x = rng.uniform(100, 300, 2000)
    u = rng.normal(0, 0.05, (2000, 3))
    E = rng.normal(0, 1e3, N + 1)
    H = rng.normal(0, 1e3, N + 1)
    b1 = _beam(x, u)
    J1 = P.push_deposit(b1, E, H, H)
    gam = np.sqrt(1 + np.sum(b1.u ** 2, axis=1))
    assert abs(J1.sum() - b1.jscale * np.sum(b1.u[:, 0] / gam)) < 1e-12 * np.abs(J1).sum()
    perm = rng.permutation(2000)
    b2 = _beam(x[perm], u[perm])
    J2 = P.push_deposit(b2, E, H, H)
    np.testing.assert_allclose(J2, J1, rtol=0, atol=1e-14 * np.abs(J1).max())
    np.testing.assert_array_equal(b2.x, b1.x[perm])


def test_absorption_parks_and_stops():
    N = 200
    b = _beam([2.3, 150.0], [[0.01, 0.0, -0.9], [0.0, 0.0, 0.0]], N=N)
    z = np.zeros(N + 1)
    J = P.push_deposit(b, z, z, z)
    assert not b.alive[0] and b.alive[1] and b.x[0] == 1.5
    assert np.abs(J).sum() > 0               # it deposits during the step it leaves
    J = P.push_deposit(b, z, z, z)
    assert np.all(J == 0) and b.x[0] == 1.5


def test_cold_beam_drifts_without_current():
    spec = ph.default_spec(N=600, steps=50, slab_start=300, slab_stop=300, probes=[60, 500], src_amp=0.0)
    m = ph.build(spec)
    x = np.linspace(250, 350, 101)
    u = np.zeros((101, 3))
    u[:, 2] = 0.1
    b = _beam(x, u, N=600, Sc=0.5, dz=spec["dz"])
    s, _ = P.run(m, b, steps=50)
    assert np.all(s.E == 0) and np.all(s.H == 0)
    beta = 0.1 / math.sqrt(1.01)
    np.testing.assert_allclose(b.x, x + 50 * 0.5 * beta, rtol=0, atol=1e-9)


def test_c_particle_step_bitwise_equal_to_numpy_oracle():
    """oracle/fdtd_cpu.c's particle step (the compiled PIC CPU baseline): one
    thread is bitwise oracle/pic.py (fields, probes, x, u, alive) on a driven
    beam with absorption at the left end; eight threads (per-thread deposit
    partials) agree to 1e-12."""
    from oracle import cpu_ref
    spec = ph.default_spec(N=900, steps=300, slab_start=600, slab_stop=700, src_amp=1e9, probes=[50, 800])
    m = ph.build(spec)
    rng = np.random.default_rng(8)
    n = 3000
    x = rng.uniform(3, 200, n)
    u = rng.normal(0, 0.02, (n, 3))
    u[:, 2] -= 0.2                          # drifting left: some are absorbed
    def beam():
        return _beam(x, u, N=900, Sc=spec["Sc"], dz=spec["dz"])
    b = beam()
    s, pr = P.run(m, b)
    assert (~b.alive).any() and b.alive.any()
    for threads, exact in ((1, True), (8, False)):
        c = cpu_ref.CpuPic(m, beam(), threads=threads)
        c.run()
        if exact:
            assert np.array_equal(c.E, s.E) and np.array_equal(c.H, s.H)
            assert np.array_equal(c.probe_out, pr)
            assert np.array_equal(c.x, b.x) and np.array_equal(c.u, b.u)
            assert np.array_equal(c.alive.astype(bool), b.alive)
        else:
            for got, want in ((c.E, s.E), (c.probe_out, pr), (c.u, b.u)):
                assert np.linalg.norm(got - want) <= 1e-12 * np.linalg.norm(want)
