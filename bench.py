"""Benchmark: config 2 (nonlinear metamaterial) on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--precision f32|f64]

One bench step = one complete simulation of the workload (N = 65,536 cells,
50,000 time steps, Lorentz+Kerr slab of 20,000 cells with per-cell eps_inf
and w0), with coefficients already resident on the device: the timed region
holds the state reset and every stencil launch.  The metric is cell-updates
per second (N * steps / time).  The roofline object prices the stencil at
the algorithmic bytes of a one-step streaming schedule (bytes/cell-update x
cells x steps per launch) over the average launch duration measured with HIP
events on the context stream; cpu_baseline times the C restatement
(oracle/fdtd_cpu.c, FP64, OpenMP over cells) on a bounded slice of the same
workload.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "py-fdtd_pic_amd"))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# Timing events: a field-only run is timed in runs of PROFILE_EVERY
# consecutive launches between two stream markers; a PIC run stamps one step
# in PROFILE_EVERY kernel by kernel (an event pair costs the host ~20 us, more
# than a launch; stamping every launch would make the timed run host-bound).
PROFILE_EVERY = 16
METRIC = "Gcell-updates/s + Gparticle-pushes/s at \u22641e-5 rel L2 vs FP64; p50/p99 sweep latency"


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--precision", default="f32", choices=["f32", "f64"])
    ap.add_argument("--sim-steps", type=int, default=50000, help="time steps per simulation")
    ap.add_argument("--K", type=int, default=0)
    ap.add_argument("--S", type=int, default=0)
    ap.add_argument("--K-slab", type=int, default=0, help="cells per lane of slab tiles (K or K/2; 0: default)")
    ap.add_argument("--persistent", default="off", choices=["on", "off"],
                    help="config 2: persistent field launches (many blocks of S steps per launch; measured "
                         "slower than one-block launches at K=2 S=48, faster at K=1 and K=4 shapes)")
    ap.add_argument("--halo", type=int, default=0,
                    help="config 2: workgroup halo exchange every H steps (0: off)")
    ap.add_argument("--cpu-seconds", type=float, default=10.0, help="target CPU baseline sample length")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-pic", action="store_true", help="skip the config 3 (PIC) measurement")
    ap.add_argument("--pic-steps", type=int, default=20000, help="time steps of the config 3 run")
    ap.add_argument("--pic-particles", type=int, default=1_000_000)
    ap.add_argument("--pic-only", action="store_true", help="print only the config 3 object")
    ap.add_argument("--pic-group", type=int, default=0, help="cells per sort/deposit wave (0: automatic)")
    ap.add_argument("--pic-K", type=int, default=1, help="cells per lane of the config 3 field tiles")
    ap.add_argument("--pic-fusion", default="on", choices=["on", "off"],
                    help="config 3: scatter in the field step's launch (on) or a launch of its own (off)")
    ap.add_argument("--workload", default="config2", choices=["config2", "config4", "config5"],
                    help="config4: batched sweep vs the per-simulation loop; config5: served trace replay")
    ap.add_argument("--c4-count", type=int, default=4096, help="config 4 instances")
    ap.add_argument("--c4-sample", type=int, default=24, help="instances timed in the per-simulation loop")
    ap.add_argument("--c4-order", default="arrival", choices=["costly", "arrival"],
                    help="batched admission order: costly jobs first, or as listed")
    ap.add_argument("--c5-requests", type=int, default=256, help="requests of the config 5 trace replayed")
    ap.add_argument("--c5-order", default="shortest", choices=["arrival", "shortest"],
                    help="config 5 admission: least cost first (default), or first come first served")
    ap.add_argument("--c5-slots", type=int, default=128,
                    help="config 5: requests resident at most (with least-cost-first admission, 128 cut the "
                         "256-request p50 from 70.3 s to 49 s at the same p99 and throughput as 512 in arrival order)")
    ap.add_argument("--no-f64", action="store_true", help="skip the FP64 config 2 object")
    ap.add_argument("--no-c4", action="store_true", help="skip the config 4 (batched sweep) object")
    ap.add_argument("--no-c5", action="store_true", help="skip the config 5 (served trace) object")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a 256 MiB device copy (FETCH_SIZE/WRITE_SIZE calibration under rocprofv3 --pmc)")
    return ap.parse_args()


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "stencil_tiles_config2_hbm_traffic.txt")
ISSUE_FILE = os.path.join(ROOT, "profiles", "stencil_tiles_config2_valu_issue.txt")
# VALU issue peak without packed FP32: 256 CUs x 4 SIMDs x 2.4 GHz, one wave64
# instruction per 4 cycles per SIMD (G wave-instructions/s)
VALU_ISSUE_PEAK = 256 * 4 * 2.4 / 4
PIC_TRAFFIC_FILE = os.path.join(ROOT, "profiles", "pic_config3_kernel_trace_and_traffic.txt")


def calibrate_copy():
    import torch
    n = 64 * 1024 * 1024            # 64 Mi floats = 256 MiB read + 256 MiB written per copy
    a = torch.ones(n, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes": n * 4, "copies": 3}), flush=True)


def measured_traffic(precision, K, S, Ks):
    """HBM bytes per stencil launch from the committed rocprofv3 --pmc summary
    (FETCH_SIZE and WRITE_SIZE per dispatch, scaled by a copy-kernel
    calibration), or None when no summary matches this tiling."""
    if not os.path.exists(TRAFFIC_FILE):
        return None
    for line in open(TRAFFIC_FILE):
        f = line.split()
        if len(f) == 6 and f[0] == "traffic" and f[1] == precision and f[2] == f"K={K}" and f[3] == f"Kslab={Ks}" \
                and f[4] == f"S={S}":
            return float(f[5])
    return None


def measured_issue(precision, K, S, Ks):
    """VALU wave-instructions per field launch from the committed rocprofv3
    --pmc SQ_INSTS_VALU summary, or None when none matches this tiling."""
    if not os.path.exists(ISSUE_FILE):
        return None
    for line in open(ISSUE_FILE):
        f = line.split()
        if len(f) == 6 and f[0] == "issue" and f[1] == precision and f[2] == f"K={K}" and f[3] == f"Kslab={Ks}" \
                and f[4] == f"S={S}":
            return float(f[5])
    return None


def measured_pic_traffic(K, fused):
    """HBM bytes per PIC step (every kernel of the step) from the committed
    config 3 rocprofv3 --pmc summary, or None when none matches."""
    if not os.path.exists(PIC_TRAFFIC_FILE):
        return None
    for line in open(PIC_TRAFFIC_FILE):
        f = line.split()
        if len(f) == 5 and f[:2] == ["traffic", "pic"] and f[2] == f"K={K}" and f[3] == f"fused={int(fused)}":
            return float(f[4])
    return None


def cpu_baseline(p, target_s):
    """C restatement (FP64, OpenMP over cells) on the same workload: whole
    simulations are repeated (the last one cut short) until about target_s
    seconds of CPU stepping have been timed."""
    from oracle import cpu_ref, physical as ph
    try:   # tune for the host that runs the baseline; fall back to the portable in-tree build
        cpu_ref.use(cpu_ref.build_lib(path="/tmp/fdtd_cpu_native.so", native=True))
    except Exception:
        cpu_ref.use(cpu_ref.build_lib())
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or cpu_ref.lib().cpu_max_threads()
    m = ph.build(p)
    probe = 200
    rate = p["N"] * probe / cpu_run(cpu_ref, m, probe, threads)
    cells = secs = 0.0
    sims = 0
    while secs < 0.95 * target_s:
        n = int(min(p["steps"], max(probe, (target_s - secs) * rate / p["N"])))
        secs += cpu_run(cpu_ref, m, n, threads)
        cells += p["N"] * n
        rate = cells / secs
        sims += 1
    return dict(value=cells / secs / 1e9, unit="Gcell-updates/s", cores=threads, kind="port",
                sample=f"config2, {sims} run(s) from step 0 ({cells / p['N']:.0f} steps of N={p['N']} cells in "
                       f"total; FP64 C, {threads} OpenMP threads), {secs:.1f} s")


def pic_bench(a):
    """Config 3: N = 65,536 grid with the config-1 slab and a 1e6-particle
    electron beam, 20,000 steps, FP32 fields and particles.  One timed run
    after one warm-up run; particles are re-uploaded (reset) outside the
    timed region."""
    from fdtd_pic_amd import configs, engine
    p = configs.config3(steps=a.pic_steps, n_particles=a.pic_particles)
    x, u, q, m, w = configs.beam_inputs(p)
    beam = engine.Beam(x, u, q, m, w, configs.DT, p["dz"], a.pic_group)
    # one cell per lane: with one step per launch the narrow tiles run the
    # field step fastest on a single simulation (K=1 4.6 us, K=2 5.4 us)
    ctx = engine.Context("f32", K=a.pic_K, S=1)
    ctx.set_pic_fusion(a.pic_fusion == "on")
    sim = ctx.add(engine.physical_spec(p), beam=beam)
    ctx.run_all()
    ctx.sync()
    sim.reset()
    ctx.sync()
    ctx.set_profile(True, every=PROFILE_EVERY)
    ctx.timing(reset=True)
    t0 = time.perf_counter()
    ctx.run_all()
    ctx.sync()
    wall = time.perf_counter() - t0
    kt = {k: ctx.kernel_timing(k) for k in ("stencil", "push", "sort", "field+sort")}
    kt = {k: v for k, v in kt.items() if v[1] > 0}
    ctx.close()
    n, T, N = p["beam"]["n"], p["steps"], p["N"]
    pushes, cells = n * T, N * T
    bpc = configs.bytes_per_cell_update(p, "f32")
    bpp = 40.0          # id int32 + frac + ux, uy, uz (FP32), read and written (one-cell beams keep no cell array)
    # per-step time of the whole timed run
    step_ms = wall * 1e3 / T
    achieved = (n * bpp + N * bpc) / (step_ms / 1e3) / 1e9
    out = {"metric": "Gparticle-pushes/s", "value": round(pushes / wall / 1e9, 3),
           "cells_value": round(cells / wall / 1e9, 3), "ms_per_run": round(wall * 1e3, 2),
           "config": {"workload": f"config3: N={N} cells, {T} steps, config-1 slab, {n} electrons "
                                  "(0.1c drift, 1e-3c thermal), FP32 + FP64 deposit sums", "K": a.pic_K, "S": 1,
                      "fused_field_scatter": a.pic_fusion == "on"},
           "kernels_avg_us": {k: round(v[0] * 1e3 / max(v[1], 1), 2) for k, v in kt.items()},
           "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(achieved / HBM_PEAK_GBS, 4), "bytes_per_push": bpp,
                        "bytes_per_cell_update": round(bpc, 2),
                        "traffic": measured_pic_traffic(a.pic_K, a.pic_fusion == "on"),
                        "traffic_per": "step (push/count + field step + scatter)"}}
    if not a.no_cpu:
        out["cpu_baseline"] = pic_cpu_baseline(p, x, u, q, m, w)
    return out


def pic_cpu_baseline(p, x, u, q, m, w, target_s=10.0):
    """The compiled particle + field step (oracle/fdtd_cpu.c: OpenMP push,
    per-thread deposit partials, OpenMP field step; FP64) on every host
    thread, over the first steps of the same workload until about target_s
    seconds have been timed."""
    from oracle import cpu_ref, pic as P, physical as ph
    try:
        cpu_ref.use(cpu_ref.build_lib(path="/tmp/fdtd_cpu_native.so", native=True))
    except Exception:
        cpu_ref.use(cpu_ref.build_lib())
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or cpu_ref.lib().cpu_max_threads()
    mm = ph.build(p)
    c = cpu_ref.CpuPic(mm, P.Beam(x, u, q, m, w, mm["dt"], p["N"], p["Sc"]), threads=threads)
    probe = 5
    secs = c.run(0, probe)
    steps = probe
    while secs < 0.95 * target_s and steps < p["steps"]:
        n = int(min(p["steps"] - steps, max(probe, (target_s - secs) / (secs / steps))))
        secs += c.run(steps, steps + n)
        steps += n
    return dict(value=len(x) * steps / secs / 1e9, unit="Gparticle-pushes/s", cores=threads, kind="port",
                sample=f"config3, first {steps} steps of {len(x)} particles with the field step, compiled FP64 C "
                       f"(oracle/fdtd_cpu.c), {threads} OpenMP threads, {secs:.1f} s")


def config4_bench(a):
    """Config 4: 4,096 two-pass sweep points (1,024 with beams) batched on one
    GPU, against the plain per-simulation loop (each job alone, one after the
    other) timed on a seeded sample of the same jobs and scaled by work."""
    from fdtd_pic_amd import configs, serve, sweep
    inst = configs.config4_instances()[:a.c4_count]
    jobs = [sweep.Job(i, p) for i, p in enumerate(inst)]
    sch = sweep.Scheduler("f32", budget_bytes=200 << 30, autotune=True)
    t0 = time.perf_counter()
    sch.run_batch(jobs, costly_first=a.c4_order == "costly")
    wall = time.perf_counter() - t0
    tiling = (sch.K, sch.S)
    host = {"admission_s": round(sch.admit_s, 3), "admission_prepare_s": round(sch.admit_prep_s, 3),
            "admission_add_s": round(sch.admit_add_s, 3), "device_wait_s": round(sch.wait_s, 3)}
    sch.close()
    cells = sum(serve.work(p)[0] for p in inst)
    pushes = sum(serve.work(p)[1] for p in inst)
    # per-simulation loop: a seeded sample from each stratum (with / without a
    # beam) run alone one after the other; each stratum's time per unit of
    # work (cell-updates + 4 x particle-pushes) scales to its total work.  A
    # beam alone gets the push/count group size best for one beam (~48
    # particles per group: enough waves to fill the GPU; 12 seeded lone beam
    # jobs took 2.46 s against 2.51 s at ~96 and 2.63 s at ~128), not the
    # larger groups the library picks by default for a full batched context
    def lone(p):
        if not p.get("beam"):
            return p
        b = p["beam"]
        ppc = b["n"] / (b["cell_hi"] - b["cell_lo"])
        q = dict(p)
        q["beam"] = dict(b, group=max(1, min(64, int(round(48.0 / ppc)))))
        return q

    import numpy as np
    rng = np.random.default_rng(44)
    w = lambda p: serve.work(p)[0] + 4 * serve.work(p)[1]
    est, alone, sample = 0.0, 0.0, []
    for has_beam in (False, True):
        idx = [i for i, p in enumerate(inst) if bool(p.get("beam")) == has_beam]
        if not idx:
            continue
        pick = sorted(rng.choice(idx, size=min(a.c4_sample // 2, len(idx)), replace=False).tolist())
        t1 = time.perf_counter()
        for i in pick:
            sweep.run_alone(lone(inst[i]), "f32")
        dt = time.perf_counter() - t1
        alone += dt
        sample += pick
        est += dt * sum(w(inst[i]) for i in idx) / sum(w(inst[i]) for i in pick)
    return {"metric": "config4 sweep: batched vs per-simulation GPU loop", "jobs": len(inst),
            "beam_jobs": sum(1 for p in inst if p.get("beam")), "batched_s": round(wall, 3),
            "gcell_updates_per_s": round(cells / wall / 1e9, 3),
            "gparticle_pushes_per_s": round(pushes / wall / 1e9, 3), "autotuned_tiling": list(tiling),
            "per_sim_loop_sample": f"{len(sample)} jobs, half with beams, seeded; scaled per stratum by work",
            "per_sim_loop_sample_s": round(alone, 3),
            "per_sim_loop_est_s": round(est, 3), "speedup_vs_per_sim_loop": round(est / wall, 2),
            "admission_order": a.c4_order, "batched_host": host,
            "spectra": "R(f), T(f) at 256 frequencies per job, FP64 DFT on the device"}


def config5_bench(a):
    """Config 5: the first requests of the seeded Poisson trace, replayed in
    real time (overload: the full trace offers ~14x what one GPU sustains)."""
    from fdtd_pic_amd import configs, serve
    arr, ps = configs.config5_requests()
    n = a.c5_requests
    srv = serve.Server("f32", budget_bytes=200 << 30, max_slots=a.c5_slots, order=a.c5_order)
    last = [time.perf_counter()]

    def log(msg):                        # completions, at most one line per 20 s
        if time.perf_counter() - last[0] >= 20.0:
            last[0] = time.perf_counter()
            progress(f"config5 {msg}")
    jobs, wall = srv.replay(arr[:n], ps[:n], log=log)
    srv.close()
    out = serve.summarize(jobs, wall)
    out.update(metric="config5 served stream (trace prefix replay)",
               trace=f"first {n} of 2000 requests, Poisson 200/s, seed 5", admission=a.c5_order,
               slots=a.c5_slots)
    return out


def cpu_run(cpu_ref, m, n, threads):
    return cpu_ref.CpuSim(m).run(0, n, threads)


def config2_line(a, precision):
    """The config 2 measurement in one precision: (bench line dict, parameters)."""
    from fdtd_pic_amd import configs, engine

    p = configs.config2(steps=a.sim_steps)
    ctx = engine.Context(precision)
    persistent = a.persistent == "on"
    K0, S0, Ks0 = (configs.TILING_CONFIG2_PERSISTENT if persistent else configs.TILING_CONFIG2)[precision]
    K = a.K or K0
    Ks = a.K_slab or (Ks0 if K == K0 else K)
    S = a.S or (S0 if K == K0 else 8 * Ks)
    ctx.set_tiling(K, S, Ks)
    ctx.set_persistent(persistent)
    ctx.set_halo_exchange(a.halo)
    sim = ctx.add(engine.physical_spec(p))
    persistent = ctx.persistent_active()
    halo = a.halo if ctx.halo_exchange_active() else 0
    for _ in range(a.warmup):
        sim.reset()
        ctx.run_all()
    ctx.sync()
    ctx.set_profile(True, every=PROFILE_EVERY)
    ctx.timing(reset=True)
    t0 = time.perf_counter()
    for _ in range(a.steps):
        sim.reset()
        ctx.run_all()
    ctx.sync()
    wall = time.perf_counter() - t0
    kern_ms, sampled = ctx.timing()
    ms_per_step = wall * 1e3 / a.steps
    cells = p["N"] * p["steps"]
    value = cells / (wall / a.steps) / 1e9
    bpc = configs.bytes_per_cell_update(p, precision)
    launch_ms = kern_ms / max(sampled, 1)                 # average launch-to-launch duration
    achieved = bpc * p["N"] * S / (launch_ms / 1e3) / 1e9 if kern_ms > 0 else 0.0
    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "Gcell-updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32+fp64-kerr" if precision == "f32" else "f64",
        "data": "synthetic (seeded config 2 parameters)",
        "config": {"workload": "config2: N=65536 cells, 50000 steps, Lorentz+Kerr slab [20000,40000)",
                   "sim_steps": p["steps"], "K": K, "K_slab": Ks, "S": S, "precision": precision,
                   "persistent": persistent, "halo_exchange": halo},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": measured_traffic(precision, K, S, Ks),
                     "algorithmic_bytes_per_launch": round(bpc * p["N"] * S),
                     "bytes_per_cell_update": round(bpc, 2),
                     "avg_launch_us": round(launch_ms * 1e3, 2), "sampled_launches": sampled,
                     "launches": a.steps * ((p["steps"] + S - 1) // S),
                     "note": "achieved = algorithmic bytes of the one-step streaming count per launch over the "
                             "launch time: an effective figure (temporal blocking keeps the state on chip); "
                             "the DRAM bytes the kernel moves are `traffic`, dram_frac their rate over the peak"},
    }
    tr = out["roofline"]["traffic"]
    if tr and kern_ms > 0:
        out["roofline"]["dram_frac"] = round(tr / (launch_ms / 1e3) / 1e9 / HBM_PEAK_GBS, 4)
    issued = measured_issue(precision, K, S, Ks)
    if issued and kern_ms > 0 and not persistent:
        rate = issued / (launch_ms / 1e3) / 1e9
        U = 64 * K - 2 * S
        out["issue"] = {"bound": "valu_issue", "achieved": round(rate, 1), "peak": VALU_ISSUE_PEAK,
                        "unit": "G wave-instructions/s", "frac": round(rate / VALU_ISSUE_PEAK, 4),
                        "wave_insts_per_launch": issued,
                        "lane_insts_per_cell_update": round(issued * 64 / (p["N"] * S), 1),
                        "useful_fraction": round(U / (64 * K), 3),
                        "useful_lane_insts_per_cell_update": round(issued * 64 / (p["N"] * S) * U / (64 * K), 1),
                        "source": "rocprofv3 --pmc SQ_INSTS_VALU (profiles/stencil_tiles_config2_valu_issue.txt) "
                                  "over the live launch time"}
    ctx.close()
    return out, p


def progress(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (--gpus {a.gpus} requested): simulations shard across GPUs "
              "as independent instances or 1-D domain slices, and that path is not built", file=sys.stderr)
        sys.exit(2)
    if a.calibrate_copy:
        calibrate_copy()
        return
    if a.pic_only:
        print(json.dumps(pic_bench(a)), flush=True)
        return
    if a.workload == "config4":
        print(json.dumps(config4_bench(a)), flush=True)
        return
    if a.workload == "config5":
        print(json.dumps(config5_bench(a)), flush=True)
        return
    out, p = config2_line(a, a.precision)
    progress(f"config2 {a.precision}: {out['value']} Gcell-updates/s")
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(p, a.cpu_seconds)
        progress(f"config2 cpu baseline: {out['cpu_baseline']['value']:.3f} Gcell-updates/s")
    if not a.no_pic:
        out["pic"] = pic_bench(a)
        progress(f"config3 pic: {out['pic']['value']} Gparticle-pushes/s")
    if not a.no_f64 and a.precision != "f64":
        f64, _ = config2_line(a, "f64")
        out["config2_f64"] = {k: f64[k] for k in ("value", "unit", "ms_per_step", "dtype", "config", "roofline")}
        progress(f"config2 f64: {f64['value']} Gcell-updates/s")
    if not a.no_c4:
        out["config4"] = config4_bench(a)
        progress(f"config4: {out['config4']['batched_s']} s batched, {out['config4']['speedup_vs_per_sim_loop']}x")
    if not a.no_c5:
        out["config5"] = config5_bench(a)
        progress(f"config5: p50 {out['config5']['p50_latency_s']} s, p99 {out['config5']['p99_latency_s']} s")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
