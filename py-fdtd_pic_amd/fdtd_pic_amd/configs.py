"""Seeded synthetic workloads (physical model parameter dicts).

lambda0 = 1.5 um, dz = lambda0/40 = 37.5 nm, Sc = 0.5, CPML width 32 on
both ends.  Every random quantity is drawn from numpy.random.default_rng
with the config number as seed, in FP64, so the oracle and the device see
identical inputs.

config1  N=2000, 10,000 steps, Lorentz slab [1000,1400) (eps_inf 1.5,
         d_eps 2, w0 2pi 150 THz, gamma 2pi 5 THz), soft Gaussian source
         (200 THz, tau 30 fs) at cell 100, probes at 60 and N-42.
config2  N=65,536, 50,000 steps, slab [20000,40000) Lorentz + Kerr
         chi3 = 1e-19 m^2/V^2, eps_inf ~ U[1,3], w0 jittered +-5% per cell,
         source amplitude 1.8e8 V/m (peak index change ~1e-3).
config3  N=65,536, 20,000 steps, config-1 slab, 1e6-particle electron beam
         over cells [30000,40000), v = 0.1c z + N(0,(1e-3 c)^2).
config4  4,096 instances, Nz log-uniform [1e3, 65,536], source centre
         U[100,300] THz, d_eps U[0.5,3], chi3 U[0,2e-19], 10,000 steps; slab
         20% of Nz from 0.5 Nz; 1,024 instances with 1e4-1e5 particles.
config5  Poisson arrivals (200/s) of config-3-style requests; see serve.py.
"""
import math

import numpy as np

C0 = 299792458.0
LAMBDA0 = 1.5e-6
DZ = LAMBDA0 / 40
SC = 0.5
DT = SC * DZ / C0
W = 32
TWO_PI = 2 * math.pi
E_CHARGE = 1.602176634e-19
M_ELECTRON = 9.1093837139e-31      # CODATA 2022


# Tiling of the config-2 loop measured best on MI355X: (K, S, K_slab) =
# cells per lane of plain tiles, steps per launch, cells per lane of tiles
# that meet the slab.
TILING_CONFIG2 = {"f32": (2, 48, 2), "f64": (2, 48, 2)}
# With persistent launches a tile's halo must lie in its neighbours'
# interiors (S <= 64K - 2S): at K = 2, S <= 42.
TILING_CONFIG2_PERSISTENT = {"f32": (2, 42, 2), "f64": (2, 42, 2)}


def base(N, steps, **kw):
    p = dict(N=int(N), steps=int(steps), Sc=SC, dz=DZ, pml_width=W, pml_order=3, pml_R0=1e-8,
             slab_start=int(N) // 2, slab_stop=int(N) // 2, eps_inf=1.5, d_eps=2.0,
             omega0=TWO_PI * 150e12, gamma=TWO_PI * 5e12, chi3=0.0, ade=True,
             src_cell=100, src_f0=200e12, src_tau=30e-15, src_t0=120e-15, src_amp=1.0,
             probes=[60, int(N) - W - 10])
    p.update(kw)
    return p


def config1():
    return base(2000, 10000, slab_start=1000, slab_stop=1400)


def config2(seed=2, steps=50000, N=65536):
    rng = np.random.default_rng(seed)
    F, R = 20000 * N // 65536, 40000 * N // 65536
    ns = R - F
    eps_inf = rng.uniform(1.0, 3.0, ns)
    omega0 = TWO_PI * 150e12 * (1.0 + rng.uniform(-0.05, 0.05, ns))
    return base(N, steps, slab_start=F, slab_stop=R, eps_inf=eps_inf, omega0=omega0, chi3=1e-19,
                src_amp=1.8e8)


def config3(seed=3, steps=20000, N=65536, n_particles=1_000_000):
    p = base(N, steps, slab_start=1000, slab_stop=1400)
    p["beam"] = dict(seed=seed, n=int(n_particles), cell_lo=30000 * N // 65536, cell_hi=40000 * N // 65536,
                     v_drift=0.1 * C0, v_th=1e-3 * C0, density=1e16)
    return p


def beam_inputs(p):
    """Seeded electron beam of a parameter dict with a "beam" entry:
    positions uniform over [cell_lo, cell_hi) (cell units), velocity
    v_drift z + N(0, v_th^2) per component, u = gamma v / c; macro-particle
    weight w = density * extent / n per unit transverse area.  Returns
    (x, u, q, m, w)."""
    bm = p["beam"]
    rng = np.random.default_rng(bm["seed"])
    n = int(bm["n"])
    x = bm["cell_lo"] + rng.uniform(0.0, bm["cell_hi"] - bm["cell_lo"], n)
    v = rng.normal(0.0, bm["v_th"], (n, 3))
    v[:, 2] += bm["v_drift"]
    beta = v / C0
    u = beta / np.sqrt(1.0 - np.sum(beta * beta, axis=1))[:, None]
    w = bm["density"] * (bm["cell_hi"] - bm["cell_lo"]) * p["dz"] / n
    return x, u, -E_CHARGE, M_ELECTRON, w


def config4_instances(seed=4, count=4096, steps=10000):
    rng = np.random.default_rng(seed)
    Nz = np.round(np.exp(rng.uniform(math.log(1e3), math.log(65536), count))).astype(int)
    fc = rng.uniform(100e12, 300e12, count)
    d_eps = rng.uniform(0.5, 3.0, count)
    chi3 = rng.uniform(0.0, 2e-19, count)
    beam_ids = set(rng.permutation(count)[:count // 4].tolist())
    nparts = rng.integers(10_000, 100_001, count)
    out = []
    for i in range(count):
        N = int(Nz[i])
        p = base(N, steps, slab_start=N // 2, slab_stop=N // 2 + N // 5, d_eps=float(d_eps[i]),
                 chi3=float(chi3[i]), src_f0=float(fc[i]), src_amp=1.8e8 if chi3[i] > 0 else 1.0,
                 src_cell=min(100, N // 10))
        p["probes"] = [max(W + 4, p["src_cell"] - 40), N - W - 10]
        if i in beam_ids:
            lo = N // 8 * 5
            p["beam"] = dict(seed=seed * 100003 + i, n=int(nparts[i]), cell_lo=lo,
                             cell_hi=min(lo + N // 8, N - W - 20), v_drift=0.1 * C0, v_th=1e-3 * C0,
                             density=1e16)
        out.append(p)
    return out


def config5_requests(seed=5, count=2000, rate=200.0):
    """Served stream: Poisson arrivals (mean gap 1/rate s), config-3-style
    requests with Nz = min(262,144, round(LogNormal(ln 8192, 1))) (at least
    2,000), Np ~ U_int[0, 1e6] electrons over cells [0.458 Nz, 0.61 Nz) (the
    config-3 beam window), T ~ U_int[5,000, 50,000] steps and the config-1
    slab at [1000, 1400).  Returns (arrival times in s, parameter dicts)."""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate, count)
    arrivals = np.cumsum(gaps) - gaps[0]
    Nz = np.minimum(262144, np.round(rng.lognormal(math.log(8192), 1.0, count))).astype(int)
    Nz = np.maximum(Nz, 2000)
    Np = rng.integers(0, 1_000_001, count)
    T = rng.integers(5000, 50_001, count)
    out = []
    for i in range(count):
        N = int(Nz[i])
        p = base(N, int(T[i]), slab_start=1000, slab_stop=1400)
        if Np[i] > 0:
            lo, hi = int(0.458 * N), int(0.61 * N)
            p["beam"] = dict(seed=seed * 100003 + i, n=int(Np[i]), cell_lo=lo, cell_hi=hi,
                             v_drift=0.1 * C0, v_th=1e-3 * C0, density=1e16)
        out.append(p)
    return arrivals, out


def bytes_per_cell_update(p, precision="f32"):
    """Algorithmic HBM bytes per cell-update of the one-step streaming schedule.

    vacuum: E, H read+write; CPML adds psi_E, psi_H; a Lorentz slab cell adds
    D, P, P_prev (and in FP32 the two Kahan words), plus per-cell eps_inf and
    w0 reads when they vary per cell.  Averaged over the grid.
    """
    w = 8 if precision == "f64" else 4
    N = p["N"]
    pml = 2 * p["pml_width"]
    ns = max(p["slab_stop"] - p["slab_start"], 0)
    vac = N - pml - ns
    b_vac = 4 * w
    b_pml = 8 * w
    b_slab = 10 * w + (0 if precision == "f64" else 4 * w)
    if np.ndim(p["eps_inf"]) or np.ndim(p["omega0"]):
        b_slab += 2 * w
    return (vac * b_vac + pml * b_pml + ns * b_slab) / N
