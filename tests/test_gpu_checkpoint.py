"""Checkpoint / resume (fdtd_state_size, fdtd_state_save, fdtd_state_load).

A simulation saved after a whole number of S-step launches and loaded into
a new simulation of the same description, in another context of the same
tiling, finishes bitwise equal to the run that was saved: fields, ADE and
CPML state, probes, snapshot rows and energy rows, and with a beam (one-cell
and multi-cell groups) the particles and J'.  The resuming context also runs
a fresh copy from step 0 beside it, which must match too (the step position
and ping-pong parity are per simulation).  A checkpoint is refused, before
anything is copied, by a simulation of another shape, precision, beam or
energy setting, and when truncated or not a checkpoint.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FIELDS = ("E", "H", "D", "P", "PSIE", "PSIH")


def _field_spec():
    from fdtd_pic_amd import engine
    from oracle import physical as ph
    p = ph.default_spec(N=700, steps=1800, slab_start=330, slab_stop=470, src_cell=60, probes=[40, 650],
                        src_t0=60e-15, src_tau=15e-15, chi3=1e-3, src_amp=2.0)
    s = engine.physical_spec(p)
    s.snap_interval, s.snap_rows = 100, 18
    return s


def _field_out(sim):
    out = {k: sim.read(k) for k in FIELDS}
    out["probes"] = sim.probes()
    out["snap"] = sim.snapshots()
    out["energy_steps"], out["energy"] = sim.energy()
    out["progress"] = sim.progress()
    return out


def _same(a, b, tag):
    for k in a:
        assert np.array_equal(np.asarray(a[k]), np.asarray(b[k])), (tag, k)


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_field_resume_bitwise(prec):
    from fdtd_pic_amd import engine
    print("field checkpoint", prec, flush=True)
    with engine.Context(prec, K=2, S=48) as ca:
        a = ca.add(_field_spec())
        a.set_energy(1, 64)
        ca.run(480)                      # ten launches
        blob = a.save_state()
        ca.run_all()
        ref = _field_out(a)
    assert ref["progress"] == 1800 and len(ref["energy"]) == 38
    with engine.Context(prec, K=2, S=48) as cb:
        fresh = cb.add(_field_spec())
        fresh.set_energy(1, 64)
        cb.run(96)                       # the resuming context is two launches in
        b = cb.add(_field_spec())
        b.set_energy(1, 64)
        b.load_state(blob)
        assert b.progress() == 480
        cb.run_all()
        _same(ref, _field_out(b), "resumed")
        # the fresh copy started two launches before the resumed one
        _same(ref, _field_out(fresh), "fresh")


@pytest.mark.parametrize("group", [1, 0])
@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_beam_resume_bitwise(prec, group):
    from fdtd_pic_amd import configs, engine
    from test_gpu_pic import _params
    p = _params("driven", steps=400, n=6000)
    x, u, q, mm, w = configs.beam_inputs(p)
    dt = p["Sc"] * p["dz"] / configs.C0
    beam = lambda: engine.Beam(x, u, q, mm, w, dt, p["dz"], group)
    print("beam checkpoint", prec, "group", group, flush=True)

    def out(sim):
        xo, uo, alive = sim.particles()
        return dict(E=sim.read("E"), H=sim.read("H"), J=sim.read("J"), probes=sim.probes(), x=xo, u=uo,
                    alive=alive, progress=sim.progress())
    with engine.Context(prec) as ca:
        a = ca.add(engine.physical_spec(p), beam=beam())
        ca.run(150)
        blob = a.save_state()
        ca.run_all()
        ref = out(a)
    with engine.Context(prec) as cb:
        fresh = cb.add(engine.physical_spec(p), beam=beam())
        cb.run(7)
        b = cb.add(engine.physical_spec(p), beam=beam())
        b.load_state(blob)
        cb.run_all()
        _same(ref, out(b), "resumed")
        _same(ref, out(fresh), "fresh")


def test_checkpoint_refused():
    from fdtd_pic_amd import _capi as capi, configs, engine
    from test_gpu_pic import _params
    with engine.Context("f64", K=2, S=48) as c:
        a = c.add(_field_spec())
        c.run(48)
        blob = a.save_state()
        other = c.add(engine.physical_spec(dict(_params("driven", N=1024, steps=200), beam=None)))
        with pytest.raises(capi.FdtdError) as e:
            other.load_state(blob)                  # another shape
        assert e.value.code == capi.FDTD_E_ARG
        same = c.add(_field_spec())
        for bad in (blob[:40], blob[:-8], b"X" + blob[1:]):
            with pytest.raises(capi.FdtdError) as e:
                same.load_state(bad)                # truncated, or not a checkpoint
            assert e.value.code == capi.FDTD_E_ARG
        with pytest.raises(capi.FdtdError) as e:
            same.set_energy(1, 8)
            same.load_state(blob)                   # energy recording differs
        assert e.value.code == capi.FDTD_E_ARG
    with engine.Context("f32", K=2, S=48) as c:
        s = c.add(_field_spec())
        with pytest.raises(capi.FdtdError) as e:
            s.load_state(blob)                      # another precision
        assert e.value.code == capi.FDTD_E_ARG
    p = _params("driven", steps=100, n=2000)
    x, u, q, mm, w = configs.beam_inputs(p)
    dt = p["Sc"] * p["dz"] / configs.C0
    with engine.Context("f64") as c:
        withb = c.add(engine.physical_spec(p), beam=engine.Beam(x, u, q, mm, w, dt, p["dz"], 0))
        c.run(10)
        bblob = withb.save_state()
        nob = c.add(engine.physical_spec(dict(p, beam=None)))
        with pytest.raises(capi.FdtdError) as e:
            nob.load_state(bblob)                   # the beam is missing
        assert e.value.code == capi.FDTD_E_ARG
