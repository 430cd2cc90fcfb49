"""Workgroup halo exchange (fdtd_set_halo_exchange) on the GPU.

The four waves of a field workgroup take adjacent windows that overlap by
2h cells and replace each other's inner halos through LDS every h steps.
Against the FP64 oracle (oracle/physical.py): relative L2 <= 1e-10 on E, H,
D, P, psi and the probes in FP64, <= 1e-5 in FP32, at several (K, S, h);
against the default tiles at the same (K, S): bitwise equal in FP64 (the
FP64 field update rounds the same way in every tiling), with snapshots and
energy rows.  A context of simulations of different lengths, one
admitted mid-run, gives each the result it gets alone, bitwise.  The mode is
declined where it does not apply (a beam, K = 1, h not a multiple of K,
h > S).
"""
import numpy as np
import pytest

from refcases import rel_l2

pytestmark = pytest.mark.gpu

FIELDS = ("E", "H", "D", "P", "PSIE", "PSIH")


def _spec(**kw):
    from oracle import physical as ph
    # the pulse (source at cell 100, half a cell per step) is inside the slab
    # [700, 1100) by the last steps
    base = dict(N=1400, steps=1800, chi3=1e-3, src_amp=2.0, slab_start=700, slab_stop=1100)
    base.update(kw)
    return ph.default_spec(**base)


def _device(spec, prec, K, S, h, energy=False, snap=False):
    from fdtd_pic_amd import engine
    with engine.Context(prec, K=K, S=S) as ctx:
        ctx.set_halo_exchange(h)
        s = engine.physical_spec(spec)
        if snap:
            s.snap_interval, s.snap_rows = 100, 15
        sim = ctx.add(s)
        if energy:
            sim.set_energy(1, 64)
        assert ctx.halo_exchange_active() == (h > 0)
        ctx.run_all()
        out = {k: sim.read(k) for k in FIELDS}
        out["probes"] = sim.probes()
        if snap:
            out["snap"] = sim.snapshots()
        if energy:
            out["energy"] = sim.energy()[1]
    return out


@pytest.mark.parametrize("K,S,h", [(2, 48, 8), (2, 32, 4), (2, 64, 16), (4, 96, 8), (4, 64, 32)])
@pytest.mark.parametrize("prec,tol", [("f64", 1e-10), ("f32", 1e-5)])
def test_halo_exchange_matches_oracle(K, S, h, prec, tol):
    from oracle import physical as ph
    spec = _spec()
    print("halo", prec, "K", K, "S", S, "h", h, flush=True)
    s, pr = ph.run(ph.build(spec))
    d = _device(spec, prec, K, S, h)
    errs = dict(E=rel_l2(d["E"], s.E), H=rel_l2(d["H"], s.H), D=rel_l2(d["D"], s.D), P=rel_l2(d["P"], s.P),
                probes=rel_l2(d["probes"], pr), psiE=rel_l2(d["PSIE"], s.psiE), psiH=rel_l2(d["PSIH"], s.psiH))
    print(errs, flush=True)
    for k, v in errs.items():
        assert v <= tol, (k, v)


@pytest.mark.parametrize("K,S,h", [(2, 48, 8), (4, 96, 8)])
def test_halo_exchange_equals_default_tiles(K, S, h):
    spec = _spec()
    print("halo vs default", K, S, h, flush=True)
    a = _device(spec, "f64", K, S, h, energy=True, snap=True)
    b = _device(spec, "f64", K, S, 0, energy=True, snap=True)
    for k in a:
        assert np.array_equal(a[k], b[k]), (k, rel_l2(a[k], b[k]))


def test_halo_exchange_mix_bitwise():
    from fdtd_pic_amd import engine
    specs = [_spec(N=1500, steps=900), _spec(N=3100, steps=1400, slab_start=1700, slab_stop=2300),
             _spec(N=900, steps=700, slab_start=450, slab_stop=850, probes=[40, 700])]

    def alone(sp):
        with engine.Context("f32", K=2, S=48) as c:
            c.set_halo_exchange(8)
            sim = c.add(engine.physical_spec(sp))
            c.run_all()
            return {k: sim.read(k) for k in ("E", "H", "D", "P")} | {"probes": sim.probes()}
    ref = [alone(sp) for sp in specs]
    with engine.Context("f32", K=2, S=48) as c:
        c.set_halo_exchange(8)
        sims = [c.add(engine.physical_spec(specs[0])), c.add(engine.physical_spec(specs[1]))]
        c.run(96)
        sims.append(c.add(engine.physical_spec(specs[2])))
        assert c.halo_exchange_active()
        c.run_all()
        got = [{k: sm.read(k) for k in ("E", "H", "D", "P")} | {"probes": sm.probes()} for sm in sims]
    for i, (r, g) in enumerate(zip(ref, got)):
        for k in r:
            assert np.array_equal(r[k], g[k]), (i, k)


def test_halo_exchange_declined():
    from fdtd_pic_amd import configs, engine
    from test_gpu_pic import _params
    cases = [(1, 16, 8), (2, 48, 3), (2, 8, 16), (4, 16, 24)]
    for K, S, h in cases:
        with engine.Context("f64", K=K, S=S) as c:
            c.set_halo_exchange(h)
            c.add(engine.physical_spec(_spec(steps=100)))
            assert not c.halo_exchange_active(), (K, S, h)
    p = _params("driven", steps=50, n=1000)
    x, u, q, mm, w = configs.beam_inputs(p)
    dt = p["Sc"] * p["dz"] / configs.C0
    with engine.Context("f64", K=2, S=48) as c:
        c.set_halo_exchange(8)
        c.add(engine.physical_spec(p), beam=engine.Beam(x, u, q, mm, w, dt, p["dz"], 0))
        assert not c.halo_exchange_active()
