"""Spectra, batched sweeps and serving on the GPU.

* The device FP64 DFT of a probe series equals numpy's (relative 1e-12).
* A batch of config-4-shaped jobs (two passes each, FP32) gives the same
  spectra as the FP64 oracle: |X32 - X64| <= 1e-4 * max|X64| at every one of
  the 256 frequencies, and |R32 - R64|, |T32 - T64| <= 1e-4 wherever the
  incident spectrum is above 10% of its peak (R and T are ratios: an X error
  of e * peak becomes e / 0.1 in them there, and outside the source band they
  divide rounding noise by almost nothing).
* Alone vs mix: each job's spectra are bitwise identical when it runs by
  itself and when it shares the contexts with other jobs (field-only and
  beam jobs), batched or served from an arrival trace.
* A served job with a beam matches the FP64 particle oracle to 1e-4.
* FP64 under the autotuner: the tiling it picks depends on timing, and the
  FP64 field update rounds the same way in every tiling, so a job's FP64
  spectra in an autotuned batch equal the job alone at the default tiling
  bitwise, and so does the job alone at each tiling the autotuner can pick.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FREQS = np.linspace(100e12, 300e12, 256)


def _small_instances(count=4, steps=4000, seed=41, beams=(1,), n_particles=3000):
    from fdtd_pic_amd import configs
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        N = int(rng.integers(900, 2400))
        chi3 = float(rng.uniform(0, 2e-19))
        p = configs.base(N, steps, slab_start=N // 2, slab_stop=N // 2 + N // 5,
                         d_eps=float(rng.uniform(0.5, 3.0)), chi3=chi3, src_f0=float(rng.uniform(150e12, 250e12)),
                         src_amp=1.8e8 if chi3 > 0 else 1.0, src_cell=min(100, N // 10))
        p["probes"] = [max(36, p["src_cell"] - 40), N - 42]
        if i in beams:
            lo = N // 8 * 5
            p["beam"] = dict(seed=1000 + i, n=n_particles, cell_lo=lo, cell_hi=min(lo + N // 8, N - 52),
                             v_drift=0.1 * configs.C0, v_th=1e-3 * configs.C0, density=1e16)
        out.append(p)
    return out


def _oracle_spectra(p):
    from fdtd_pic_amd import configs, sweep
    from oracle import physical as ph, pic as P
    X = []
    for with_slab in (False, True):
        q = sweep.pass_params(p, with_slab)
        m = ph.build(q)
        if q.get("beam"):
            x, u, qq, mm, w = configs.beam_inputs(q)
            _, probes = P.run(m, P.Beam(x, u, qq, mm, w, m["dt"], q["N"], q["Sc"]))
        else:
            _, probes = ph.run(m)
        X.append(ph.dft(probes, m["dt"], FREQS))
    (X0, X1) = X
    return X0[0], X1[0], X0[1], X1[1]


def test_device_dft_matches_numpy():
    from fdtd_pic_amd import configs, engine
    p = configs.base(1500, 3000, slab_start=750, slab_stop=1000)
    dt = p["Sc"] * p["dz"] / configs.C0
    for prec in ("f64", "f32"):
        with engine.Context(prec) as ctx:
            sim = ctx.add(engine.physical_spec(p))
            ctx.run_all()
            pr = sim.probes()
            X = ctx.spectra([sim, sim], 1, FREQS, dt)
        ref = engine.dft(pr[1], dt, FREQS)[0]
        err = np.max(np.abs(X[0] - ref)) / np.max(np.abs(ref))
        print(prec, "dft err", err, flush=True)
        assert err < 1e-12 and np.array_equal(X[0], X[1])


@pytest.fixture(scope="module")
def batch_jobs():
    from fdtd_pic_amd import sweep
    ps = _small_instances()
    sch = sweep.Scheduler("f32", K=2, S=48, freqs=FREQS)
    jobs = [sweep.Job(i, p) for i, p in enumerate(ps)]
    sch.run_batch(jobs)
    sch.close()
    return ps, jobs


def test_batch_spectra_vs_oracle(batch_jobs):
    ps, jobs = batch_jobs
    for p, j in zip(ps, jobs):
        X0r, X1r, X0t, X1t = _oracle_spectra(p)
        scale_r, scale_t = np.abs(X0r).max(), np.abs(X0t).max()
        ex = max(np.abs(j.X[0] - X0r).max() / scale_r, np.abs(j.X[1] - X1r).max() / scale_r,
                 np.abs(j.X[2] - X0t).max() / scale_t, np.abs(j.X[3] - X1t).max() / scale_t)
        band_r = np.abs(X0r) > 0.1 * scale_r
        band_t = np.abs(X0t) > 0.1 * scale_t
        R64 = np.abs(X1r - X0r) / np.abs(X0r)
        T64 = np.abs(X1t) / np.abs(X0t)
        eR = np.abs(j.R - R64)[band_r].max()
        eT = np.abs(j.T - T64)[band_t].max()
        print(f"job {j.id} N={p['N']} beam={bool(p.get('beam'))} X {ex:.2e} R {eR:.2e} T {eT:.2e}", flush=True)
        assert ex <= 1e-4 and eR <= 1e-4 and eT <= 1e-4


def test_alone_equals_batch_bitwise(batch_jobs):
    from fdtd_pic_amd import sweep
    ps, jobs = batch_jobs
    for p, j in zip(ps, jobs):
        a = sweep.run_alone(p, "f32", K=2, S=48, freqs=FREQS)
        for k in range(4):
            assert np.array_equal(a.X[k], j.X[k]), (j.id, k)


def test_busy_reports_queued_work():
    """fdtd_busy: set while queued launches run, clear after fdtd_sync."""
    from fdtd_pic_amd import configs, engine
    p = configs.config2(steps=9600)   # 200 launches of ~15 us each: the device lags the host
    with engine.Context("f32", K=2, S=48) as ctx:
        ctx.add(engine.physical_spec(p))
        ctx.run(9600)
        seen = ctx.busy()
        ctx.sync()
        assert seen and not ctx.busy()


def test_admission_while_running_bitwise():
    """Jobs admitted while slices of others run (upload stream, first chunk of
    one job) give the same spectra, bitwise, as each job alone."""
    from fdtd_pic_amd import sweep
    ps = _small_instances(count=6, steps=2000, seed=13, beams=(1, 4), n_particles=2000)
    sch = sweep.Scheduler("f32", K=2, S=48, freqs=FREQS, max_slots=3)
    jobs = [sweep.Job(i, p) for i, p in enumerate(ps)]
    try:
        sch.run_batch(jobs, first_chunk=1)
    finally:
        sch.close()
    assert all(j.X is not None for j in jobs)
    for p, j in zip(ps, jobs):
        a = sweep.run_alone(p, "f32", K=2, S=48, freqs=FREQS)
        for k in range(4):
            assert np.array_equal(a.X[k], j.X[k]), (j.id, k)


def test_served_trace_alone_vs_mix():
    from fdtd_pic_amd import serve, sweep
    ps = _small_instances(count=5, steps=3000, seed=7, beams=(0, 3))
    arrivals = np.array([0.0, 0.01, 0.02, 0.2, 0.21])
    srv = serve.Server("f32", max_slots=3, freqs=FREQS)
    jobs, wall = srv.replay(arrivals, ps)
    srv.close()
    s = serve.summarize(jobs, wall)
    print(s, flush=True)
    assert all(j.finished is not None and j.finished >= j.arrival for j in jobs)
    for p, j in zip(ps, jobs):
        a = sweep.run_alone(p, "f32", freqs=FREQS)
        for k in range(4):
            assert np.array_equal(a.X[k], j.X[k]), (j.id, k)
    # a beam job of the trace against the FP64 particle oracle
    X0r, X1r, X0t, X1t = _oracle_spectra(ps[0])
    ex = np.abs(jobs[0].X[1] - X1r).max() / np.abs(X0r).max()
    print("served beam job vs oracle", ex, flush=True)
    assert ex <= 1e-4


def test_served_least_cost_first():
    """order="shortest": requests waiting together are admitted in order of
    cost (cell-updates + 4 x particle-pushes), here two at a time; each still
    gets the spectra it gets alone, bitwise."""
    from fdtd_pic_amd import serve, sweep
    ps = _small_instances(count=5, steps=2500, seed=19, beams=(1, 2), n_particles=2000)
    srv = serve.Server("f32", max_slots=2, freqs=FREQS, order="shortest")
    jobs, wall = srv.replay(np.zeros(len(ps)), ps)
    srv.close()
    by_admission = [j.id for j in sorted(jobs, key=lambda j: j.admitted)]
    by_cost = sorted(range(len(ps)), key=lambda i: serve.cost(ps[i]))
    print("admitted", by_admission, "by cost", by_cost, flush=True)
    assert by_admission == by_cost
    for p, j in zip(ps, jobs):
        a = sweep.run_alone(p, "f32", freqs=FREQS)
        for k in range(4):
            assert np.array_equal(a.X[k], j.X[k]), (j.id, k)


def test_nonfinite_job_flagged_and_isolated():
    """A job whose field turns non-finite mid-run (NaN written into one cell
    of its slab pass's E from the host after the first slice) finishes
    flagged failed, with non-finite spectra; the jobs sharing its contexts
    are not flagged and their spectra equal their lone runs bitwise."""
    from fdtd_pic_amd import sweep
    ps = _small_instances(count=3, steps=2400, seed=23, beams=(1,), n_particles=2000)
    sch = sweep.Scheduler("f32", K=2, S=48, freqs=FREQS)
    jobs = [sweep.Job(i, p) for i, p in enumerate(ps)]
    try:
        for j in jobs:
            sch.admit(j)
        sch.step()
        sim = jobs[0].sims[1]
        E = sim.read("E")
        E[len(E) // 2] = np.nan
        sim.write("E", E)
        while sch.active:
            sch.step()
            sch.collect()
    finally:
        sch.close()
    print("failed", [j.failed for j in jobs], flush=True)
    assert jobs[0].failed and not np.isfinite(jobs[0].X[1]).all()
    assert not jobs[1].failed and not jobs[2].failed
    for p, j in zip(ps[1:], jobs[1:]):
        a = sweep.run_alone(p, "f32", K=2, S=48, freqs=FREQS)
        for k in range(4):
            assert np.array_equal(a.X[k], j.X[k]), (j.id, k)


def test_jobs_of_different_dt_finishing_together():
    """Two jobs with different cell sizes (so different dt) that finish in the
    same slice each get the DFT at their own dt: bitwise their spectra alone."""
    from fdtd_pic_amd import configs, sweep
    ps = []
    for k, dz in enumerate((configs.DZ, 1.25 * configs.DZ)):
        p = configs.base(1200, 2400, slab_start=600, slab_stop=800, dz=dz, src_cell=80, probes=[40, 1150])
        p["src_f0"] = 200e12 / (1.0 + 0.25 * k)
        ps.append(p)
    sch = sweep.Scheduler("f32", K=2, S=48, freqs=FREQS)
    jobs = [sweep.Job(i, p) for i, p in enumerate(ps)]
    try:
        sch.run_batch(jobs)
    finally:
        sch.close()
    assert jobs[0].finished == jobs[1].finished
    for p, j in zip(ps, jobs):
        a = sweep.run_alone(p, "f32", K=2, S=48, freqs=FREQS)
        for k in range(4):
            assert np.array_equal(a.X[k], j.X[k]), (j.id, k)


def _fp64_same(a, b, tag):
    for k in range(4):
        inc = np.abs(a.X[0 if k < 2 else 2]).max()
        assert inc > 0, (tag, k)
        err = np.abs(a.X[k] - b.X[k]).max() / inc
        print(f"{tag} X{k} {err:.2e}", flush=True)
        assert np.array_equal(a.X[k], b.X[k]), (tag, k, err)


def test_fp64_autotuned_batch_vs_alone():
    from fdtd_pic_amd import sweep
    ps = _small_instances(count=3, steps=6000, seed=77, beams=())
    sch = sweep.Scheduler("f64", autotune=True, freqs=FREQS)
    jobs = [sweep.Job(i, p) for i, p in enumerate(ps)]
    try:
        sch.run_batch(jobs)
        picked = (sch.K, sch.S)
    finally:
        sch.close()
    print("autotuned tiling", picked, flush=True)
    for p, j in zip(ps, jobs):
        a = sweep.run_alone(p, "f64", freqs=FREQS)
        _fp64_same(a, j, f"job {j.id} batch")
        same = sweep.run_alone(p, "f64", K=picked[0], S=picked[1], freqs=FREQS)
        for k in range(4):
            assert np.array_equal(same.X[k], j.X[k]), (j.id, k)
    # every tiling the autotuner can pick, for one job
    a0 = sweep.run_alone(ps[0], "f64", freqs=FREQS)
    for K, S in sweep.TILINGS:
        _fp64_same(a0, sweep.run_alone(ps[0], "f64", K=K, S=S, freqs=FREQS), f"tiling {K}x{S}")
