"""Host engine over libfdtdpic.so: contexts, simulation descriptors, runs.

`Context` owns one HIP stream and any number of simulations that advance
together (one simulation for the drop-in integrators, thousands for a
batched sweep).  `SimSpec` is the Python form of fdtd_sim_desc_t: it keeps
the numpy arrays alive while the struct points at them.
"""
import ctypes as C
import math

import numpy as np

from . import _capi as capi
from ._capi import check, dptr, as_f64

C0 = 299792458.0


class SimSpec:
    """Python-side fdtd_sim_desc_t with owned arrays."""

    _ARRAYS = ("s_ke", "s_kh", "s_bE", "s_cE", "s_CbE", "s_bH", "s_cH", "s_C2H",
               "sl_A", "sl_B", "sl_C", "sl_e", "srcE", "srcH")

    def __init__(self, N, steps):
        self.N, self.steps = int(N), int(steps)
        self.ke0 = self.kh0 = 0.0
        self.Ls, self.Rs = 0, self.N + 1
        self.F = self.R = 0
        self.slab_kind, self.ade = capi.SLAB_NONE, 0
        self.keS = self.khS = self.kD = self.jD = 0.0
        self.kerr_a = self.kerr_b = self.kerr_c = 0.0
        self.kerr_e0 = self.kerr_lin = self.kerr_chi3 = 0.0
        self.src_cell, self.tfsf = 1, 0
        self.probes = []          # (cell, t0, t1)
        self.snap_interval = self.snap_rows = 0
        self.has_current = 0
        for a in self._ARRAYS:
            setattr(self, a, None)

    # -- strips from full-length per-cell coefficient arrays -------------
    def set_cells(self, ke, kh, bE, cE, CbE, bH, cH, C2H, ke0, kh0, diel=None):
        """Split full-length (N+1) coefficient arrays into interior scalars + two strips.

        The interior is the longest run of cells whose coefficients equal the
        interior values (or keS/khS inside a DIEL slab) and carry no CPML.
        """
        n = self.N + 1
        arrs = [as_f64(a) for a in (ke, kh, bE, cE, CbE, bH, cH, C2H)]
        for a in arrs:
            if a.shape != (n,):
                raise ValueError("coefficient arrays must have length N+1")
        ke, kh = arrs[0], arrs[1]
        expect_ke = np.full(n, float(ke0))
        expect_kh = np.full(n, float(kh0))
        if diel is not None:
            F, R, keS, khS = diel
            expect_ke[F:R] = keS
            expect_kh[F:R] = khS
        plain = (ke == expect_ke) & (kh == expect_kh)
        for a in arrs[2:]:
            plain &= (a == 0.0)
        # longest run of plain cells (the first one on ties)
        edges = np.diff(np.concatenate(([0], plain.astype(np.int8), [0])))
        starts, ends = np.flatnonzero(edges == 1), np.flatnonzero(edges == -1)
        best = (0, 0)
        if len(starts):
            k = int(np.argmax(ends - starts))
            best = (int(starts[k]), int(ends[k]))
        Ls, Rs = best if best[1] > best[0] else (n, n)
        self.Ls, self.Rs = int(Ls), int(Rs)
        self.ke0, self.kh0 = float(ke0), float(kh0)
        idx = np.r_[0:Ls, Rs:n]
        for name, a in zip(("s_ke", "s_kh", "s_bE", "s_cE", "s_CbE", "s_bH", "s_cH", "s_C2H"), arrs):
            setattr(self, name, np.ascontiguousarray(a[idx]))

    def to_struct(self):
        d = capi.SimDesc()
        for f in ("N", "steps", "Ls", "Rs", "F", "R", "slab_kind", "ade", "src_cell", "tfsf",
                  "snap_interval", "snap_rows", "has_current"):
            setattr(d, f, int(getattr(self, f)))
        for f in ("ke0", "kh0", "keS", "khS", "kD", "jD", "kerr_a", "kerr_b", "kerr_c",
                  "kerr_e0", "kerr_lin", "kerr_chi3"):
            setattr(d, f, float(getattr(self, f)))
        self._keep = []
        for a in self._ARRAYS:
            v = getattr(self, a)
            if v is not None:
                v = as_f64(v)
                self._keep.append(v)
                setattr(d, a, dptr(v))
        if len(self.probes) > capi.MAX_PROBES:
            raise ValueError(f"at most {capi.MAX_PROBES} probes")
        d.nprobe = len(self.probes)
        for k, (c, t0, t1) in enumerate(self.probes):
            d.probe_cell[k], d.probe_t0[k], d.probe_t1[k] = int(c), int(t0), int(t1)
        return d


class Sim:
    def __init__(self, ctx, sid, spec):
        self.ctx, self.id, self.spec = ctx, sid, spec
        self.beam_n = 0

    def read(self, field):
        s = self.spec
        n = {"E": s.N + 1, "H": s.N + 1, "PSIE": s.N + 1, "PSIH": s.N + 1, "J": s.N + 1}.get(
            field, max(s.R - s.F, 0))
        out = np.zeros(n)
        check(self.ctx.lib.fdtd_read_field(self.ctx.h, self.id, capi.FIELD[field], dptr(out), n))
        return out

    def write(self, field, values):
        v = as_f64(values)
        check(self.ctx.lib.fdtd_write_field(self.ctx.h, self.id, capi.FIELD[field], dptr(v), v.size))

    def probes(self):
        s = self.spec
        out = np.zeros((len(s.probes), s.steps))
        check(self.ctx.lib.fdtd_read_probes(self.ctx.h, self.id, dptr(out), out.size))
        return out

    def snapshots(self):
        s = self.spec
        out = np.zeros((s.snap_rows, s.N + 1))
        check(self.ctx.lib.fdtd_read_snapshots(self.ctx.h, self.id, dptr(out), out.size))
        return out

    def set_energy(self, every=1, rows=1024):
        """Record the FP64 field energy after every `every`-th launch (0: off)."""
        check(self.ctx.lib.fdtd_set_energy(self.ctx.h, self.id, int(every), int(rows)))
        self.energy_rows = int(rows)

    def energy(self):
        """(steps, W): the step index after which each energy value was taken, and W."""
        cap = getattr(self, "energy_rows", 0)
        vals = np.zeros(max(cap, 1))
        steps = np.zeros(max(cap, 1), dtype=np.int64)
        n = C.c_int64()
        check(self.ctx.lib.fdtd_read_energy(self.ctx.h, self.id, dptr(vals),
                                            steps.ctypes.data_as(C.POINTER(C.c_int64)), cap, C.byref(n)))
        k = min(n.value, cap)
        return steps[:k], vals[:k]

    def progress(self):
        v = C.c_int64()
        check(self.ctx.lib.fdtd_sim_progress(self.ctx.h, self.id, C.byref(v)))
        return v.value

    def particles(self):
        """(x, u, alive) of the attached beam in input order: x in cell units,
        u = gamma v / c as n x 3, alive as bool."""
        n = self.beam_n
        x, u = np.zeros(n), np.zeros((n, 3))
        alive = np.zeros(n, dtype=np.int8)
        check(self.ctx.lib.pic_read(self.ctx.h, self.id, dptr(x), dptr(u),
                                    alive.ctypes.data_as(C.POINTER(C.c_int8)), n))
        return x, u, alive.astype(bool)

    def reset(self):
        check(self.ctx.lib.fdtd_reset_sim(self.ctx.h, self.id))

    def save_state(self):
        """Checkpoint of this simulation (and its beam) after every queued
        launch, as bytes (fdtd_state_save)."""
        n = C.c_int64()
        check(self.ctx.lib.fdtd_state_size(self.ctx.h, self.id, C.byref(n)))
        buf = C.create_string_buffer(n.value)
        check(self.ctx.lib.fdtd_state_save(self.ctx.h, self.id, buf, n.value))
        return buf.raw

    def load_state(self, blob):
        """Continue from a checkpoint of a simulation built from the same spec
        (same precision, beam and energy setting; fdtd_state_load)."""
        buf = C.create_string_buffer(bytes(blob), len(blob))
        check(self.ctx.lib.fdtd_state_load(self.ctx.h, self.id, buf, len(blob)))

    def remove(self):
        check(self.ctx.lib.fdtd_remove_sim(self.ctx.h, self.id))


class Context:
    """One HIP stream plus the simulations it advances (fdtd_ctx)."""

    def __init__(self, precision="f64", K=None, S=None):
        self.lib = capi.load()
        self.precision = precision
        h = C.c_void_p()
        check(self.lib.fdtd_ctx_create(capi.F64 if precision == "f64" else capi.F32, C.byref(h)))
        self.h = h
        if K is not None:
            check(self.lib.fdtd_set_tiling(self.h, int(K), int(S or 0)))

    def set_tiling(self, K, S=0, K_slab=None):
        """K cells per lane, S steps per launch; K_slab = K/2 makes tiles that
        meet a Lorentz/Kerr slab narrower than plain ones."""
        check(self.lib.fdtd_set_tiling_mixed(self.h, int(K), int(K if K_slab is None else K_slab), int(S)))

    def set_pic_fusion(self, on=True):
        """Run the particle scatter in the field step's launch when that
        launch is small (default on; bitwise the same results either way)."""
        check(self.lib.fdtd_set_pic_fusion(self.h, 1 if on else 0))

    def set_pic_lds_rows(self, rows=0):
        """LDS rows per cell of the one-cell particle kernel (0: chosen per
        launch from the beams' densities; bitwise the same results)."""
        check(self.lib.fdtd_set_pic_lds_rows(self.h, int(rows)))

    def set_persistent(self, on=True):
        """Persistent field launches: many blocks of S steps per launch when
        every tile fits on the device at once (bitwise the same results)."""
        check(self.lib.fdtd_set_persistent(self.h, 1 if on else 0))

    def persistent_active(self):
        """True when the next run() uses persistent launches."""
        v = C.c_int()
        check(self.lib.fdtd_persistent_active(self.h, C.byref(v)))
        return bool(v.value)

    def set_halo_exchange(self, h=8):
        """Workgroup halo exchange every h steps between the four adjacent
        windows of a field workgroup (0: off); declined where it does not apply."""
        check(self.lib.fdtd_set_halo_exchange(self.h, int(h)))

    def halo_exchange_active(self):
        """True when the next run() uses the workgroup halo exchange."""
        v = C.c_int()
        check(self.lib.fdtd_halo_exchange_active(self.h, C.byref(v)))
        return bool(v.value)

    def add(self, spec, beam=None):
        """Add a simulation; beam (optional) is a Beam whose current couples to it."""
        sid = C.c_int()
        if beam is not None:
            spec.has_current = 1
        d = spec.to_struct()
        check(self.lib.fdtd_add_sim(self.h, C.byref(d), C.byref(sid)))
        sim = Sim(self, sid.value, spec)
        if beam is not None:
            x, u = as_f64(beam.x), as_f64(beam.u).reshape(-1, 3)
            bd = capi.BeamDesc(n=len(x), q=beam.q, m=beam.m, w=beam.w, dt=beam.dt, dz=beam.dz,
                               x=dptr(x), u=dptr(u), group=int(beam.group))
            check(self.lib.pic_attach(self.h, sim.id, C.byref(bd)))
            sim.beam_n = len(x)
        return sim

    def run(self, nsteps):
        check(self.lib.fdtd_run(self.h, int(nsteps)))

    def run_all(self):
        check(self.lib.fdtd_run_all(self.h))

    def sync(self):
        check(self.lib.fdtd_sync(self.h))

    def busy(self):
        """True while launches queued by run() are still executing."""
        b = C.c_int()
        check(self.lib.fdtd_busy(self.h, C.byref(b)))
        return bool(b.value)

    def set_profile(self, on=True, every=1):
        """Timing events (off: on=False): without a beam, runs of `every`
        consecutive field launches between two stream markers; with a beam,
        every `every`-th step's kernels stamped one by one."""
        check(self.lib.fdtd_set_profile(self.h, int(every) if on else 0))

    def timing(self, reset=False):
        """(ms, launches) of the sampled field launches; ms / launches is the
        average launch duration."""
        ms, n = C.c_double(), C.c_int64()
        check(self.lib.fdtd_timing(self.h, C.byref(ms), C.byref(n), 1 if reset else 0))
        return ms.value, n.value

    def kernel_timing(self, kind):
        """(ms, launches) of one kernel kind: "stencil", "push", "sort", or
        "field+sort" (the fused field and scatter launch)."""
        k = {"stencil": 0, "push": 1, "sort": 2, "field+sort": 3}[kind]
        ms, n = C.c_double(), C.c_int64()
        check(self.lib.fdtd_kernel_timing(self.h, k, C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def spectra(self, sims, probe, freqs, dt):
        """Complex FP64 DFT (len(sims) x len(freqs)) of probe `probe` of each sim, on the device."""
        ids = np.ascontiguousarray([s.id for s in sims], dtype=np.int32)
        f = as_f64(freqs)
        out = np.zeros((len(ids), len(f), 2))
        check(self.lib.fdtd_probe_spectra(self.h, len(ids), ids.ctypes.data_as(C.POINTER(C.c_int)), int(probe),
                                          dptr(f), len(f), float(dt), dptr(out)))
        return out[..., 0] + 1j * out[..., 1]

    def device_bytes(self):
        v = C.c_int64()
        check(self.lib.fdtd_device_bytes(self.h, C.byref(v)))
        return v.value

    def close(self):
        if self.h:
            self.lib.fdtd_ctx_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Beam:
    """Macro-particle beam: positions x (cell units), u = gamma v / c (n x 3),
    charge q and mass m of one physical particle, weight w per unit
    transverse area (m^-2), the grid's dt and dz, and the cells per
    sort/deposit wave (0: chosen from the density)."""

    def __init__(self, x, u, q, m, w, dt, dz, group=0):
        self.x, self.u, self.q, self.m, self.w = x, u, float(q), float(m), float(w)
        self.dt, self.dz, self.group = float(dt), float(dz), int(group)


# ---------------------------------------------------------------------------
# Physical model descriptors (configs 1-5)
# ---------------------------------------------------------------------------
def pml_coefficients(N, W, Sc, order=3, R0=1e-8):
    """CPML b, c on E nodes (n dz) and H nodes ((n+1/2) dz); kappa = 1, alpha = 0.

    sigma dt/eps0 = Sc (m+1)(-ln R0)/(2W) * depth^m.
    """
    n = np.arange(N + 1, dtype=np.float64)
    s0 = Sc * (order + 1) * (-math.log(R0)) / (2.0 * W)
    # depth 0 (b = exp(-0) = 1 exactly, c = 0) off the strips: evaluate the
    # strip cells only
    idx = np.flatnonzero((n < W) | (n >= N - W))
    m = n[idx]
    de = np.where(m < W, (W - m) / W, (m - (N - W)) / W)
    dh = np.where(m < W, (W - m - 0.5) / W, np.minimum((m + 0.5 - (N - W)) / W, 1.0))
    bE = np.ones(N + 1)
    bH = np.ones(N + 1)
    bE[idx] = np.exp(-s0 * de ** order)
    bH[idx] = np.exp(-s0 * dh ** order)
    in_e = np.zeros(N + 1, dtype=bool)
    in_h = np.zeros(N + 1, dtype=bool)
    in_e[idx] = de > 0
    in_h[idx] = dh > 0
    return bE, bE - 1.0, bH, bH - 1.0, in_e, in_h


def lorentz_coefficients(dt, omega0, d_eps, gamma):
    """ADE P' = A P' + B P'm + C E (BaseFDTD11.py:620-626, eps0 wp^2 -> d_eps w0^2)."""
    D = 1.0 / dt ** 2 + gamma / (2.0 * dt)
    A = (2.0 / dt ** 2 - omega0 ** 2) / D
    B = (gamma / (2.0 * dt) - 1.0 / dt ** 2) / D
    Cc = (d_eps * omega0 ** 2) / D
    return A, np.broadcast_to(B, np.shape(A)).astype(np.float64), Cc


def gaussian_source(steps, dt, amp, f0, tau, t0):
    t = np.arange(steps, dtype=np.float64) * dt - t0
    return amp * np.exp(-(t / tau) ** 2) * np.sin(2.0 * math.pi * f0 * t)


def physical_spec(p, with_slab=True, probe_window=None):
    """SimSpec for the physical model from a parameter dict (same keys as the
    configs: N, steps, Sc, dz, pml_width, pml_order, pml_R0, slab_start,
    slab_stop, eps_inf, d_eps, omega0, gamma, chi3, ade, src_*, probes)."""
    N, T, Sc = int(p["N"]), int(p["steps"]), float(p["Sc"])
    W = int(p["pml_width"])
    dt = Sc * p["dz"] / C0
    s = SimSpec(N, T)
    bE, cE, bH, cH, me, mh = pml_coefficients(N, W, Sc, p.get("pml_order", 3), p.get("pml_R0", 1e-8))
    me[0] = me[N] = False
    mh[N] = False
    ke = np.full(N + 1, Sc)
    ke[0] = ke[N] = 0.0
    kh = np.full(N + 1, Sc)
    kh[N] = 0.0
    z = np.zeros(N + 1)
    s.set_cells(ke, kh, np.where(me, bE, 0.0), np.where(me, cE, 0.0), np.where(me, -Sc, 0.0),
                np.where(mh, bH, 0.0), np.where(mh, cH, 0.0), np.where(mh, Sc, 0.0), Sc, Sc)
    if with_slab and p["slab_stop"] > p["slab_start"]:
        F, R = int(p["slab_start"]), int(p["slab_stop"])
        ns = R - F
        cell = lambda v: np.full(ns, float(v)) if np.ndim(v) == 0 else as_f64(v)
        s.F, s.R = F, R
        s.slab_kind = capi.SLAB_KERR_E
        s.ade = 1 if p.get("ade", True) else 0
        A, B, Cc = lorentz_coefficients(dt, cell(p["omega0"]), cell(p["d_eps"]), float(p["gamma"]))
        s.sl_A, s.sl_B, s.sl_C = A, B, Cc
        s.sl_e = cell(p["eps_inf"])
        s.kerr_chi3 = float(p.get("chi3", 0.0))
        s.kD, s.jD = Sc, 1.0
    s.src_cell = int(p["src_cell"])
    s.srcE = gaussian_source(T, dt, p["src_amp"], p["src_f0"], p["src_tau"], p["src_t0"])
    w0, w1 = probe_window or (0, T - 1)
    s.probes = [(int(c), w0, w1) for c in p.get("probes", [])]
    del z
    return s


def dft(x, dt, freqs):
    """FP64 DFT X(f) = sum_t x_t exp(-2 pi i f t dt) dt over the last axis."""
    x = np.atleast_2d(x)
    t = np.arange(x.shape[-1]) * dt
    out = np.empty((x.shape[0], len(freqs)), dtype=np.complex128)
    for i, f in enumerate(freqs):
        out[:, i] = x @ np.exp(-2j * math.pi * f * t) * dt
    return out
