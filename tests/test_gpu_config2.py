"""Config 2 at full size (N=65,536, 50,000 steps, Lorentz+Kerr slab with
per-cell eps_inf and w0, chi3 = 1e-19, 1.8e8 V/m source) against the FP64
C oracle (oracle/fdtd_cpu.c, bitwise equal to oracle/physical.py).

FP32 (Kahan-compensated D and P; u = D - P, the small-z Kerr correction and
E in FP32, the closed-form Kerr root in FP64): relative L2 <= 1e-5 on the
fields E, H and the probe series.  At 200 THz most slab cells have eps < 0, so
at the end of the run the slab holds only an evanescent remnant (~3e-5 of the
field scale): D' and P' (V/m, same units as E) are compared relative to the
field scale ||E||, because their own norm magnifies the field-level FP32
error by ~1e4.  FP64: every array <= 1e-10 relative to its own norm.  The oracle's C run is shared
by both tests (module fixture, ~10 s on 16 threads).
"""
import numpy as np
import pytest

from refcases import rel_l2

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def oracle_run():
    from fdtd_pic_amd import configs
    from oracle import cpu_ref, physical as ph
    p = configs.config2()
    m = ph.build(p)
    print("oracle: config2 C FP64 run", flush=True)
    sim, secs = cpu_ref.run_model(m, threads=cpu_ref.lib().cpu_max_threads())
    print(f"oracle done in {secs:.1f} s", flush=True)
    return p, sim


def _device(p, prec):
    from fdtd_pic_amd import configs, engine
    K, S, Ks = configs.TILING_CONFIG2[prec]          # the tiling bench.py measures
    with engine.Context(prec) as ctx:
        ctx.set_tiling(K, S, Ks)
        sim = ctx.add(engine.physical_spec(p))
        ctx.run_all()
        return {k: sim.read(k) for k in ("E", "H", "D", "P")} | {"probes": sim.probes()}


@pytest.mark.parametrize("prec,tol", [("f32", 1e-5), ("f64", 1e-10)])
def test_config2_full_size(oracle_run, prec, tol):
    p, o = oracle_run
    d = _device(p, prec)
    scale = np.linalg.norm(o.E)
    errs = dict(E=rel_l2(d["E"], o.E), H=rel_l2(d["H"], o.H), probes=rel_l2(d["probes"], o.probe_out))
    if prec == "f64":
        errs.update(D=rel_l2(d["D"], o.D), P=rel_l2(d["P"], o.P))
    else:
        errs.update(D_vs_field=np.linalg.norm(d["D"] - o.D) / scale, P_vs_field=np.linalg.norm(d["P"] - o.P) / scale)
    print(prec, {k: f"{v:.2e}" for k, v in errs.items()}, "max|E|", float(np.abs(o.E).max()),
          "max|E| slab", float(np.abs(o.E[p["slab_start"]:p["slab_stop"]]).max()), flush=True)
    assert np.abs(o.E).max() > 1e8 and np.abs(o.D).max() > 1e3
    for k, v in errs.items():
        assert v <= tol, (prec, k, v)
