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


class Server:
    def __init__(self, precision="f32", budget_bytes=64 << 30, max_slots=512, freqs=None):
        self.sch = sweep.Scheduler(precision, budget_bytes=budget_bytes, max_slots=max_slots, freqs=freqs)

    def replay(self, arrivals, params, time_scale=1.0, log=None):
        """Replay a trace in wall-clock time (arrival t is released at t * time_scale)."""
        jobs = [sweep.Job(i, p, a * time_scale) for i, (a, p) in enumerate(zip(arrivals, params))]
        waiting = []
        nxt = 0
        t0 = time.perf_counter()
        sch = self.sch
        while nxt < len(jobs) or waiting or sch.active:
            now = time.perf_counter() - t0
            while nxt < len(jobs) and jobs[nxt].arrival <= now:
                waiting.append(jobs[nxt])
                nxt += 1
            while waiting and (sch.fits(waiting[0]) or not sch.active):
                sch.admit(waiting.pop(0), now=time.perf_counter() - t0)
            if not sch.active:
                if nxt < len(jobs):
                    time.sleep(max(0.0, jobs[nxt].arrival - (time.perf_counter() - t0)))
                continue
            sch.step()
            fin = sch.collect(now=time.perf_counter() - t0)
            if log and fin:
                log(f"t={time.perf_counter() - t0:.3f}s done {[j.id for j in fin]} active {len(sch.active)} "
                    f"waiting {len(waiting)}")
        wall = time.perf_counter() - t0
        return jobs, wall

    def close(self):
        self.sch.close()


def summarize(jobs, wall):
    lat = np.array([j.finished - j.arrival for j in jobs])
    cells = sum(work(j.p)[0] for j in jobs)
    pushes = sum(work(j.p)[1] for j in jobs)
    return dict(requests=len(jobs), wall_s=round(wall, 3),
                p50_latency_s=round(float(np.percentile(lat, 50)), 4),
                p99_latency_s=round(float(np.percentile(lat, 99)), 4),
                requests_per_s=round(len(jobs) / wall, 3),
                gcell_updates_per_s=round(cells / wall / 1e9, 3),
                gparticle_pushes_per_s=round(pushes / wall / 1e9, 3))
