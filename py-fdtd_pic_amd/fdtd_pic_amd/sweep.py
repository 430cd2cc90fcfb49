"""Many independent simulations on one GPU: batched sweeps and served streams.

The reference runs one simulation at a time: LoopedSim
(MasterController.py:530-569) loops an input-frequency sweep over full
Controller runs, each a two-pass reflection measurement
(IntegratorLinLor1D, Solver_Engine.py:275-371: pass 0 without the slab,
pass 1 with it) followed by the spectral stage (TransformHandler.RefTester,
TransformHandler.py:34-78; MasterController.results :472-505).  Here a Job
is one such point (both passes) and a Scheduler keeps many jobs resident:

* field-only passes share one context (temporally blocked, K cells per lane,
  S steps per launch); passes with a beam share a second context on its own
  stream (one step per launch, particle kernels batched over every beam);
* a job is admitted when its device bytes fit the HBM budget and a slot is
  free; its state stays on the device until both passes finish, then its
  spectra are computed on the device (FP64 DFT of the probe series) and the
  simulations are freed;
* contexts advance in slices that are multiples of S, so a job's launch
  boundaries relative to its own start do not depend on what else runs:
  the result of a job is bitwise the same alone and in a mix;
* an autotuner picks (K, S) for the field context on the admitted set
  (timed trial launches, then every simulation is reset).

Spectra (two-pass protocol): R(f) = |X1_r - X0_r| / |X0_r| at the reflection
probe, T(f) = |X1_t| / |X0_t| at the transmission probe.
"""
import math
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np

from . import configs, engine, reference_loop

DEFAULT_FREQS = np.linspace(100e12, 300e12, 256)
TILINGS = [(1, 16), (1, 24), (2, 32), (2, 40), (2, 48), (4, 64), (4, 96)]
_TUNE_CACHE = {}


def pass_params(p, with_slab):
    q = dict(p)
    if not with_slab:
        q["slab_stop"] = q["slab_start"]
    return q


def estimate_bytes(p, precision="f32"):
    """Device bytes of one job (both passes), from the library's block layout."""
    w = 8 if precision == "f64" else 4
    N, T = p["N"], p["steps"]
    ns = max(p["slab_stop"] - p["slab_start"], 0)
    strip = 2 * p["pml_width"] + 4
    per_pass = w * (8 * strip + 4 * ns + 2 * T + 2 * (4 * (N + 1) + 5 * ns + 2 * strip)
                    + len(p.get("probes", [])) * T) + 64 * 256
    slab_pass = per_pass
    vac_pass = per_pass - w * (4 * ns + 10 * ns)
    total = slab_pass + vac_pass
    if p.get("beam"):
        n = p["beam"]["n"]
        # (cell, frac, ux, uy, uz, id, start) x 2, the rank record, per-cell counts and sums
        beam = 2 * (4 * n + 4 * w * n + 4 * n + 4 * (N + 2)) + 4 * n + 40 * (N + 2) + 16 * 256
        total += 2 * (beam + w * (N + 1))
    return int(total)


def job_dt(p):
    """Time step of a job's simulations (s)."""
    return p["Sc"] * p["dz"] / configs.C0


def job_cost(p):
    """Device bytes moved by both passes of a job, in the one-step streaming
    count: ~16 B per cell-update and ~100 B per particle-push (the sort's
    records included)."""
    n = p["beam"]["n"] if p.get("beam") else 0
    return 2 * p["steps"] * (16 * p["N"] + 100 * n)


class Job:
    """One sweep point: two passes of a parameter dict (see configs)."""

    def __init__(self, jid, p, arrival=0.0):
        self.id, self.p, self.arrival = jid, p, float(arrival)
        self.sims = []           # [vacuum pass, slab pass]
        self.ctx = None
        self.admitted = self.finished = None
        self.R = self.T = None
        self.X = None            # (X0_r, X1_r, X0_t, X1_t)
        self.failed = False      # its spectra came out non-finite (a field diverged or was corrupted)
        self.bytes = 0
        self.prepared = None     # Future of prepare(self) while queued

    @property
    def has_beam(self):
        return bool(self.p.get("beam"))


class SpecJob(Job):
    """A job given as device SimSpecs, e.g. the passes of a reference
    integrator call (reference_loop.reference_spec, planned by the drop-in
    Solver_Engine).  It is admitted like any job (field context, HBM budget)
    and on completion every pass's arrays are read back
    (reference_loop.read_pass) into `outputs`, in spec order, instead of
    spectra."""

    def __init__(self, jid, specs, arrival=0.0):
        super().__init__(jid, None, arrival)
        self.specs = list(specs)
        self.outputs = None

    @property
    def has_beam(self):
        return False


def spec_bytes(spec, precision="f64"):
    """Device bytes of one simulation built from a SimSpec (the library's block layout)."""
    w = 8 if precision == "f64" else 4
    N, T = spec.N, spec.steps
    ns = max(spec.R - spec.F, 0)
    strip = spec.Ls + (N + 1 - spec.Rs)
    return int(w * (8 * max(strip, 1) + 4 * max(ns, 1) + 2 * T
                    + 2 * (4 * (N + 1) + 5 * max(ns, 1) + 2 * max(strip, 1))
                    + max(len(spec.probes), 1) * T + max(spec.snap_rows, 1) * (N + 1)) + 64 * 256)


def job_bytes(job, precision):
    if isinstance(job, SpecJob):
        return sum(spec_bytes(s, precision) for s in job.specs)
    return estimate_bytes(job.p, precision)


def prepare(job):
    """Host inputs of a job (the beam's particles, both passes' descriptors):
    pure NumPy, so worker threads can build them ahead of admission."""
    beam_in = configs.beam_inputs(job.p) if job.has_beam else None
    specs = [engine.physical_spec(pass_params(job.p, with_slab)) for with_slab in (False, True)]
    return beam_in, specs


class Scheduler:
    def __init__(self, precision="f32", budget_bytes=64 << 30, max_slots=4096, freqs=None, K=None, S=None,
                 autotune=False, slice_launches=8, beam_slice_steps=64, prep_workers=8, prep_ahead=64):
        self.precision = precision
        self.budget = int(budget_bytes)
        self.max_slots = int(max_slots)
        self.freqs = DEFAULT_FREQS if freqs is None else np.asarray(freqs, dtype=np.float64)
        self.field = engine.Context(precision)
        self.beam = engine.Context(precision, K=2, S=1)
        self.K, self.S = K or 2, S or 48
        self.field.set_tiling(self.K, self.S)
        self.autotune_enabled = autotune
        self.slice_launches = int(slice_launches)
        self.beam_slice_steps = int(beam_slice_steps)
        self.active, self.done = [], []
        self.used = 0
        self.tuned = False
        # host seconds spent admitting jobs (of which: waiting for prepared
        # inputs, and in the library's add calls) and blocked waiting for the device
        self.admit_s = self.wait_s = 0.0
        self.admit_prep_s = self.admit_add_s = 0.0
        # jobs prepared ahead of admission by worker threads
        self.prep_workers, self.prep_ahead = int(prep_workers), int(prep_ahead)

    # -- admission -------------------------------------------------------
    def fits(self, job):
        return len(self.active) < self.max_slots and self.used + job_bytes(job, self.precision) <= self.budget

    def admit(self, job, now=None):
        t_start = time.perf_counter()
        if isinstance(job, SpecJob):
            job.bytes = job_bytes(job, self.precision)
            job.sims = [self.field.add(spec) for spec in job.specs]
            job.ctx = self.field
            job.admitted = time.perf_counter() if now is None else now
            self.used += job.bytes
            self.active.append(job)
            self.admit_s += time.perf_counter() - t_start
            return
        ctx = self.beam if job.has_beam else self.field
        job.bytes = estimate_bytes(job.p, self.precision)
        fut = getattr(job, "prepared", None)
        beam_in, specs = fut.result() if fut is not None else prepare(job)
        job.prepared = None
        t_add = time.perf_counter()
        self.admit_prep_s += t_add - t_start
        beam = None
        if job.has_beam:
            x, u, q, m, w = beam_in
            # cells per push/count group: the beam's own choice, else the library's
            beam = engine.Beam(x, u, q, m, w, job_dt(job.p), job.p["dz"],
                               int(job.p["beam"].get("group", 0)))
        for spec in specs:
            job.sims.append(ctx.add(spec, beam=beam))
        self.admit_add_s += time.perf_counter() - t_add
        job.ctx = ctx
        job.admitted = time.perf_counter() if now is None else now
        self.used += job.bytes
        self.active.append(job)
        self.admit_s += time.perf_counter() - t_start

    # -- tuning ----------------------------------------------------------
    def autotune(self, trial_launches=4):
        """Time each tiling on the admitted field passes, keep the fastest, reset them."""
        field_jobs = [j for j in self.active if not j.has_beam]
        if not field_jobs:
            return self.K, self.S
        cells = sum(s.spec.N for j in field_jobs for s in j.sims)
        key = (self.precision, int(math.log2(max(cells, 1))))
        if key in _TUNE_CACHE:
            best = _TUNE_CACHE[key]
        else:
            rates = {}
            for K, S in TILINGS:
                self.field.set_tiling(K, S)
                self.field.run(S)                     # warm-up launch (tile list rebuild)
                self.field.sync()
                t0 = time.perf_counter()
                self.field.run(S * trial_launches)
                self.field.sync()
                rates[(K, S)] = S * trial_launches / (time.perf_counter() - t0)
            best = max(rates, key=rates.get)
            _TUNE_CACHE[key] = best
            for j in field_jobs:
                for s in j.sims:
                    s.reset()
        self.K, self.S = best
        self.field.set_tiling(*best)
        self.tuned = True
        return best

    # -- stepping --------------------------------------------------------
    def launch(self):
        """Queue one slice on each context that has active jobs (asynchronous)."""
        if any(not j.has_beam for j in self.active):
            self.field.run(self.S * self.slice_launches)
        if any(j.has_beam for j in self.active):
            self.beam.run(self.beam_slice_steps)

    def wait(self):
        t_start = time.perf_counter()
        self.field.sync()
        self.beam.sync()
        self.wait_s += time.perf_counter() - t_start

    def busy(self):
        return self.field.busy() or self.beam.busy()

    def step(self):
        """Advance both contexts by one slice (each on its own stream)."""
        self.launch()
        self.wait()

    def collect(self, now=None):
        """Finish jobs whose two passes are complete: spectra on the device, then free."""
        fin = [j for j in self.active if all(s.progress() >= s.spec.steps for s in j.sims)]
        if not fin:
            return []
        for j in fin:
            if isinstance(j, SpecJob):
                j.outputs = [reference_loop.read_pass(sim, spec) for sim, spec in zip(j.sims, j.specs)]
        for ctx in (self.field, self.beam):
            # one device DFT per context and time step: jobs of a sweep over
            # Sc or dz that finish together each get their own dt
            groups = {}
            for j in fin:
                if j.ctx is ctx and not isinstance(j, SpecJob):
                    groups.setdefault(job_dt(j.p), []).append(j)
            for dt, mine in groups.items():
                sims = [s for j in mine for s in j.sims]
                Xr = ctx.spectra(sims, 0, self.freqs, dt)
                Xt = ctx.spectra(sims, 1, self.freqs, dt)
                for k, j in enumerate(mine):
                    X0r, X1r, X0t, X1t = Xr[2 * k], Xr[2 * k + 1], Xt[2 * k], Xt[2 * k + 1]
                    j.X = (X0r, X1r, X0t, X1t)
                    # a diverged or corrupted field shows as non-finite spectra: the
                    # job is flagged (the reference's loop would sys.exit), and the
                    # jobs beside it, each in simulations of its own, are unaffected
                    j.failed = not all(np.isfinite(x).all() for x in j.X)
                    with np.errstate(divide="ignore", invalid="ignore"):   # no incident power at f: nan
                        j.R = np.abs(X1r - X0r) / np.abs(X0r)
                        j.T = np.abs(X1t) / np.abs(X0t)
        t = time.perf_counter() if now is None else now
        for j in fin:
            j.finished = t
            for s in j.sims:
                s.remove()
            j.sims = []
            self.used -= j.bytes
            self.active.remove(j)
            self.done.append(j)
        return fin

    def run_batch(self, jobs, first_chunk=64, costly_first=False):
        """Run jobs to completion, admitted in list order, or costly ones
        (beams) first with costly_first (longest-processing-time order; on
        config 4 it measured slower: every job then becomes resident at once,
        58.6 s against 54.5 s).  Admission (parameter construction, uploads,
        particle sort) is host work: after a first chunk it proceeds while the
        device runs the queued slice, one job at least per slice, as long as
        the HBM budget allows; finished jobs free their room at slice ends.
        The next prep_ahead jobs' particles and descriptors are built on
        prep_workers threads meanwhile (config 4: admission was 27.7 s of
        host time in a 35.6 s batch, the device idle for most of it)."""
        pending = sorted(jobs, key=lambda j: -job_cost(j.p)) if costly_first else list(jobs)
        pool = ThreadPoolExecutor(self.prep_workers) if self.prep_workers > 1 and len(pending) > 1 else None
        try:
            return self._run(pending, pool, first_chunk)
        finally:
            if pool is not None:
                pool.shutdown(wait=True, cancel_futures=True)

    def _run(self, pending, pool, first_chunk):
        def admissible():
            return pending and (self.fits(pending[0]) or not self.active)

        def admit_next():
            # keep the next prep_ahead jobs in preparation on the worker threads
            if pool is not None:
                for job in pending[:self.prep_ahead]:
                    if getattr(job, "prepared", None) is None and not isinstance(job, SpecJob):
                        job.prepared = pool.submit(prepare, job)
            self.admit(pending.pop(0))

        n = 0
        while admissible() and n < first_chunk:
            admit_next()
            n += 1
        while pending or self.active:
            # tune once field jobs are resident (no slice is in flight here)
            if self.autotune_enabled and not self.tuned and any(not j.has_beam for j in self.active):
                self.autotune()
            floor = 1 if self.active else first_chunk     # idle device: admit a chunk before stepping
            self.launch()
            n = 0
            while admissible() and (n < floor or self.busy()):
                admit_next()
                n += 1
            self.wait()
            self.collect()
        return self.done

    def close(self):
        self.field.close()
        self.beam.close()


def run_alone(p, precision="f32", K=None, S=None, freqs=None):
    """One job by itself (the plain per-simulation loop)."""
    sch = Scheduler(precision, K=K, S=S, freqs=freqs)
    try:
        job = Job(0, p)
        sch.run_batch([job])
        return job
    finally:
        sch.close()
