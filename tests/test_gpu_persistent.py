"""Persistent field launches (fdtd_set_persistent) against one-block launches.

A persistent launch runs many blocks of S steps with every tile resident;
between blocks a tile exchanges only its halo cells with its neighbours
through per-tile flags.  The arithmetic per cell and step is the one-block
launch's, so at the same (K, S) the results must be bitwise equal: fields,
Kahan compensations (through the compensated reads), D, P, probes.  Cases:
config 2's shape (N = 65,536, Lorentz + Kerr slab with per-cell parameters)
for 2,117 steps (50 blocks of 42 and 17 steps left for a one-block launch),
FP32 and FP64; three simulations of different lengths in one context; and
the cases it must decline (a beam attached, S > 64K - 2S)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FIELDS = ("E", "H", "D", "P", "PM", "PSIE", "PSIH")


def _run(specs, prec, persistent, K, S, steps=None):
    from fdtd_pic_amd import engine
    with engine.Context(prec) as ctx:
        ctx.set_tiling(K, S)
        ctx.set_persistent(persistent)
        sims = [ctx.add(s) for s in specs]
        active = ctx.persistent_active()
        if steps is None:
            ctx.run_all()
        else:
            ctx.run(steps)
        ctx.sync()
        out = [{f: sim.read(f) for f in FIELDS} | {"probes": sim.probes()} for sim in sims]
    return active, out


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_persistent_bitwise_config2_shape(prec):
    from fdtd_pic_amd import configs, engine
    p = configs.config2(steps=2117)
    spec = lambda: engine.physical_spec(p)
    on, a = _run([spec()], prec, True, 2, 42)
    off, b = _run([spec()], prec, False, 2, 42)
    assert on and not off
    assert np.abs(a[0]["E"]).max() > 0
    for f in FIELDS + ("probes",):
        assert np.array_equal(a[0][f], b[0][f]), f


def test_persistent_bitwise_several_simulations():
    from fdtd_pic_amd import configs, engine
    ps = [configs.base(N, T, slab_start=N // 2, slab_stop=N // 2 + N // 5, chi3=1e-19 * k, src_amp=1.8e8 if k else 1.0,
                       src_cell=80, probes=[40, N - 50])
          for k, (N, T) in enumerate([(3000, 1900), (12000, 2600), (7000, 2300)])]
    on, a = _run([engine.physical_spec(p) for p in ps], "f32", True, 2, 40, steps=1000)
    off, b = _run([engine.physical_spec(p) for p in ps], "f32", False, 2, 40, steps=1000)
    assert on and not off
    for x, y in zip(a, b):
        for f in FIELDS + ("probes",):
            assert np.array_equal(x[f], y[f]), f
    # and to the end (each simulation stops at its own step count inside a launch)
    on, a = _run([engine.physical_spec(p) for p in ps], "f32", True, 2, 40)
    _, b = _run([engine.physical_spec(p) for p in ps], "f32", False, 2, 40)
    for x, y in zip(a, b):
        for f in FIELDS + ("probes",):
            assert np.array_equal(x[f], y[f]), f


def test_persistent_declined_where_it_cannot_run():
    from fdtd_pic_amd import configs, engine
    p = configs.base(4000, 200, slab_start=2000, slab_stop=2400)
    with engine.Context("f32") as ctx:
        ctx.set_tiling(2, 48)                    # S > 64K - 2S: halos reach past the neighbours
        ctx.set_persistent(True)
        ctx.add(engine.physical_spec(p))
        assert not ctx.persistent_active()
        ctx.set_tiling(2, 40)
        assert ctx.persistent_active()
    q = dict(p)
    q["beam"] = dict(seed=3, n=2000, cell_lo=1000, cell_hi=1200, v_drift=0.1 * configs.C0, v_th=1e-3 * configs.C0,
                     density=1e16)
    x, u, qq, m, w = configs.beam_inputs(q)
    with engine.Context("f32", K=2, S=40) as ctx:
        ctx.set_persistent(True)
        ctx.add(engine.physical_spec(q), beam=engine.Beam(x, u, qq, m, w, configs.DT, configs.DZ))
        assert not ctx.persistent_active()
        ctx.run(50)
        ctx.sync()
