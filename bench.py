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


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--precision", default="f32", choices=["f32", "f64"])
    ap.add_argument("--sim-steps", type=int, default=50000, help="time steps per simulation")
    ap.add_argument("--K", type=int, default=0)
    ap.add_argument("--S", type=int, default=0)
    ap.add_argument("--cpu-seconds", type=float, default=10.0, help="target CPU baseline sample length")
    ap.add_argument("--no-cpu", action="store_true")
    return ap.parse_args()


def cpu_baseline(p, target_s):
    """C restatement (FP64, OpenMP) on the first steps of the same workload."""
    from oracle import cpu_ref, physical as ph
    try:   # tune for the host that runs the baseline; fall back to the portable in-tree build
        cpu_ref.use(cpu_ref.build_lib(path="/tmp/fdtd_cpu_native.so", native=True))
    except Exception:
        cpu_ref.use(cpu_ref.build_lib())
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or cpu_ref.lib().cpu_max_threads()
    m = ph.build(p)
    sim = cpu_ref.CpuSim(m)
    probe = 500
    secs = sim.run(0, probe, threads)
    n = int(min(p["steps"] - probe, max(probe, target_s / max(secs, 1e-9) * probe)))
    secs2 = sim.run(probe, probe + n, threads)
    rate = p["N"] * n / secs2
    return dict(value=rate / 1e9, unit="Gcell-updates/s", cores=threads, kind="port",
                sample=f"config2 steps {probe}..{probe + n} of {p['steps']} (FP64 C, {threads} OpenMP threads), "
                       f"{secs2:.1f} s")


def main():
    a = parse()
    if a.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (see DESIGN.md, sharding)",
              file=sys.stderr)
        sys.exit(2)
    from fdtd_pic_amd import configs, engine

    p = configs.config2(steps=a.sim_steps)
    ctx = engine.Context(a.precision)
    K = a.K or (4 if a.precision == "f32" else 2)
    ctx.set_tiling(K, a.S)
    sim = ctx.add(engine.physical_spec(p))
    for _ in range(a.warmup):
        sim.reset()
        ctx.run_all()
    ctx.sync()
    ctx.set_profile(True)
    ctx.timing(reset=True)
    t0 = time.perf_counter()
    for _ in range(a.steps):
        sim.reset()
        ctx.run_all()
    ctx.sync()
    wall = time.perf_counter() - t0
    kern_ms, launches = ctx.timing()
    ms_per_step = wall * 1e3 / a.steps
    cells = p["N"] * p["steps"]
    value = cells / (wall / a.steps) / 1e9
    bpc = configs.bytes_per_cell_update(p, a.precision)
    alg_bytes_total = bpc * cells * a.steps
    achieved = alg_bytes_total / (kern_ms / 1e3) / 1e9 if kern_ms > 0 else 0.0
    out = {
        "metric": "Gcell-updates/s (config 2, rel L2 <= 1e-5 vs FP64 oracle)",
        "value": round(value, 3),
        "unit": "Gcell-updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32+fp64-kerr" if a.precision == "f32" else "f64",
        "data": "synthetic (seeded config 2 parameters)",
        "config": {"workload": "config2: N=65536 cells, 50000 steps, Lorentz+Kerr slab [20000,40000)",
                   "sim_steps": p["steps"], "K": K, "S": ctx_S(K, a.S), "precision": a.precision},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "bytes_per_cell_update": round(bpc, 2),
                     "avg_launch_us": round(kern_ms * 1e3 / max(launches, 1), 2), "launches": launches},
    }
    ctx.close()
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(p, a.cpu_seconds)
    print(json.dumps(out), flush=True)


def ctx_S(K, S):
    return S or 8 * K


if __name__ == "__main__":
    main()
