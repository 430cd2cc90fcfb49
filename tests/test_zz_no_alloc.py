"""fdtd_probe_spectra_noalloc (the no-allocation variant of fdtd_probe_spectra)
called through the C ABI: bitwise equal to fdtd_probe_spectra on two calls
in a row, the second with fewer series and frequencies, in both precisions;
sizes beyond its limits are refused before any device work."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _call(lib, h, fn, sims, probe, freqs, dt):
    from fdtd_pic_amd import _capi
    ids = np.ascontiguousarray([s.id for s in sims], dtype=np.int32)
    f = np.ascontiguousarray(freqs, dtype=np.float64)
    out = np.full((len(ids), len(f), 2), np.nan)
    _capi.check(getattr(lib, fn)(h, len(ids), ids.ctypes.data_as(C.POINTER(C.c_int)), int(probe), _capi.dptr(f),
                                 len(f), float(dt), _capi.dptr(out)))
    return out


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_noalloc_spectra_equal_allocating_entry_point(prec):
    from fdtd_pic_amd import configs, engine
    with engine.Context(prec) as ctx:
        sims = []
        for k, (N, T) in enumerate([(700, 1500), (900, 1800), (600, 1200)]):
            p = configs.base(N, T, slab_start=N // 2, slab_stop=N // 2 + 100, src_cell=60, probes=[40, N - 50])
            p["src_amp"] = 1.0 + k
            sims.append(ctx.add(engine.physical_spec(p)))
        ctx.run_all()
        dt = configs.DT
        f1 = np.linspace(100e12, 300e12, 256)
        a1 = _call(ctx.lib, ctx.h, "fdtd_probe_spectra", sims, 1, f1, dt)
        b1 = _call(ctx.lib, ctx.h, "fdtd_probe_spectra_noalloc", sims, 1, f1, dt)
        assert np.isfinite(b1).all() and np.abs(b1).max() > 0
        assert np.array_equal(a1, b1)
        # second call, smaller: fewer series and frequencies, the other probe
        f2 = np.linspace(150e12, 250e12, 37)
        a2 = _call(ctx.lib, ctx.h, "fdtd_probe_spectra", sims[1:], 0, f2, dt)
        b2 = _call(ctx.lib, ctx.h, "fdtd_probe_spectra_noalloc", sims[1:], 0, f2, dt)
        assert b2.shape == (2, 37, 2) and np.abs(b2).max() > 0
        assert np.array_equal(a2, b2)


def test_noalloc_spectra_refuses_sizes_beyond_limits():
    from fdtd_pic_amd import _capi, configs, engine
    with engine.Context("f32") as ctx:
        p = configs.base(600, 200, src_cell=60, probes=[40])
        sim = ctx.add(engine.physical_spec(p))
        ctx.run_all()
        ids = np.zeros(1, dtype=np.int32) + sim.id
        f = np.linspace(1e14, 3e14, 1025)
        out = np.zeros((1, 1025, 2))
        rc = ctx.lib.fdtd_probe_spectra_noalloc(ctx.h, 1, ids.ctypes.data_as(C.POINTER(C.c_int)), 0, _capi.dptr(f),
                                                1025, configs.DT, _capi.dptr(out))
        assert rc == _capi.FDTD_E_ARG and b"exceeds its limits" in ctx.lib.fdtd_last_error()
