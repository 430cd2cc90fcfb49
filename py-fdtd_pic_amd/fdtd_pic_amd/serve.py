"""Serving a stream of sweep requests on one GPU (continuous batching).

Requests arrive over time (a seeded Poisson trace, configs.config5_requests);
each is a two-pass reflection job (sweep.Job).  The server admits arrived
requests in order while their device bytes fit the HBM budget and a slot is
free, keeps every admitted job's fields and particles on the device, advances
all of them together slice by slice (field-only passes in the temporally
blocked context, beam passes in the one-step context on a second stream), and
completes a job as soon as its passes finish: spectra on the device, memory
freed, the next waiting request admitted.  Latency is completion time minus
arrival time (queueing included); throughput counts completed work.

The reference has no serving layer; its sweep driver is the blocking loop
LoopedSim (MasterController.py:530-569).
"""
import time

import numpy as np

from . import sweep


def work(p):
    """(cell-updates, particle-pushes) of one request, both passes."""
    cells = 2 * (p["N"] + 1) * p["steps"]
    pushes = 2 * p["beam"]["n"] * p["steps"] if p.get("beam") else 0
    return cells, pushes


def cost(p):
    """Relative device cost of a request: cell-updates + 4 x particle-pushes."""
    cells, pushes = work(p)
    return cells + 4 * pushes


class Server:
    """order: "arrival" admits waiting requests first come, first served;
    "shortest" admits the waiting request of least cost first (ties in
    arrival order)."""

    def __init__(self, precision="f32", budget_bytes=64 << 30, max_slots=512, freqs=None, prep_workers=8,
                 prep_ahead=64, order="arrival"):
        if order not in ("arrival", "shortest"):
            raise ValueError("order must be 'arrival' or 'shortest'")
        self.sch = sweep.Scheduler(precision, budget_bytes=budget_bytes, max_slots=max_slots, freqs=freqs)
        self.prep_workers, self.prep_ahead = int(prep_workers), int(prep_ahead)
        self.order = order

    def replay(self, arrivals, params, time_scale=1.0, log=None):
        """Replay a trace in wall-clock time (arrival t is released at t * time_scale).

        Each loop queues one slice on the contexts that hold jobs, then, while
        the device runs it (fdtd_busy), admits the requests that have arrived
        and fit, releasing new arrivals as the clock passes them; it then waits
        for the slice and completes the jobs whose passes are done.  The next
        prep_ahead waiting requests' inputs (particles, descriptors) are built
        on worker threads meanwhile.  With nothing resident it admits what has
        arrived, or sleeps until the next arrival."""
        from concurrent.futures import ThreadPoolExecutor
        jobs = [sweep.Job(i, p, a * time_scale) for i, (a, p) in enumerate(zip(arrivals, params))]
        sch = self.sch
        pool = ThreadPoolExecutor(self.prep_workers) if self.prep_workers > 1 else None
        waiting = []
        state = {"next": 0}
        t0 = time.perf_counter()
        clock = lambda: time.perf_counter() - t0

        def release():
            now = clock()
            grew = False
            while state["next"] < len(jobs) and jobs[state["next"]].arrival <= now:
                waiting.append(jobs[state["next"]])
                state["next"] += 1
                grew = True
            if grew and self.order == "shortest":
                waiting.sort(key=lambda j: cost(j.p))          # stable: ties stay in arrival order
            if pool is not None:
                for job in waiting[:self.prep_ahead]:
                    if job.prepared is None:
                        job.prepared = pool.submit(sweep.prepare, job)

        def admit_one():
            sch.admit(waiting.pop(0), now=clock())

        try:
            while state["next"] < len(jobs) or waiting or sch.active:
                release()
                if not sch.active:
                    if not waiting:
                        time.sleep(max(0.0, jobs[state["next"]].arrival - clock()))
                        continue
                    admit_one()                     # an idle device takes work at once
                    while waiting and sch.fits(waiting[0]):
                        admit_one()
                sch.launch()
                # admit while the slice runs: arrivals are released as the clock passes them
                while True:
                    release()
                    if not (waiting and sch.fits(waiting[0]) and sch.busy()):
                        break
                    admit_one()
                sch.wait()
                fin = sch.collect(now=clock())
                if log and fin:
                    log(f"t={clock():.3f}s done {[j.id for j in fin]} active {len(sch.active)} "
                        f"waiting {len(waiting)}")
        finally:
            if pool is not None:
                pool.shutdown(wait=True, cancel_futures=True)
        return jobs, clock()

    def close(self):
        self.sch.close()


def summarize(jobs, wall):
    lat = np.array([j.finished - j.arrival for j in jobs])
    cells = sum(work(j.p)[0] for j in jobs)
    pushes = sum(work(j.p)[1] for j in jobs)
    return dict(requests=len(jobs), failed=sum(1 for j in jobs if getattr(j, "failed", False)), wall_s=round(wall, 3),
                p50_latency_s=round(float(np.percentile(lat, 50)), 4),
                p99_latency_s=round(float(np.percentile(lat, 99)), 4),
                requests_per_s=round(len(jobs) / wall, 3),
                gcell_updates_per_s=round(cells / wall / 1e9, 3),
                gparticle_pushes_per_s=round(pushes / wall / 1e9, 3))
