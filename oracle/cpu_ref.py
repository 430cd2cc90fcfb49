"""Oracle: ctypes loader for fdtd_cpu.c.  TEST INFRASTRUCTURE ONLY.

`build_lib()` compiles oracle/fdtd_cpu.c with gcc -O3 -march=native
-ffp-contract=off -fopenmp into oracle/libfdtd_cpu.so; `run_model()` runs a
model dict from oracle.physical.build on it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libfdtd_cpu.so")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)


class PhysT(C.Structure):
    _fields_ = [("N", C.c_int32), ("steps", C.c_int32), ("Sc", C.c_double)] + \
        [(n, C.c_int32) for n in ("e_l0", "e_l1", "e_r0", "e_r1", "h_l0", "h_l1", "h_r0", "h_r1")] + \
        [(n, _dp) for n in ("bE", "cE", "bH", "cH")] + \
        [("F", C.c_int32), ("R", C.c_int32)] + \
        [(n, _dp) for n in ("A", "B", "C", "eps")] + \
        [("chi3", C.c_double), ("src_cell", C.c_int32), ("src", _dp), ("nprobe", C.c_int32),
         ("probes", _ip), ("probe_out", _dp), ("J", _dp)]


class StateT(C.Structure):
    _fields_ = [(n, _dp) for n in ("E", "H", "psiE", "psiH", "D", "P", "Pm")]


class BeamT(C.Structure):
    _fields_ = [("n", C.c_int64), ("x", _dp), ("u", _dp), ("alive", C.POINTER(C.c_int8)), ("a", C.c_double),
                ("jscale", C.c_double), ("Hprev", _dp), ("J", _dp), ("B1", _dp), ("B2", _dp), ("m0", _ip),
                ("w0", _dp), ("w1", _dp)]


def build_lib(force=False, path=None, native=False):
    """Compile fdtd_cpu.c.  The in-tree copy is portable (x86-64-v3); a
    `native=True` copy tuned for the host that runs the baseline can be built
    elsewhere (bench.py builds one in /tmp on the GPU box)."""
    src = os.path.join(HERE, "fdtd_cpu.c")
    out = path or LIB
    if force or not os.path.exists(out) or os.path.getmtime(out) < os.path.getmtime(src):
        arch = "-march=native" if native else "-march=x86-64-v3"
        subprocess.check_call(["gcc", "-O3", arch, "-ffp-contract=off", "-fopenmp",
                               "-fPIC", "-shared", "-o", out + ".tmp", src, "-lm"])
        os.replace(out + ".tmp", out)
    return out


_lib = None


def use(path):
    """Load the oracle library from `path` instead of the in-tree copy."""
    global _lib
    _lib = None
    return lib(path)


def lib(path=None):
    global _lib
    if _lib is None:
        p = path or LIB
        if not os.path.exists(p):
            build_lib()
        _lib = C.CDLL(p)
        _lib.cpu_phys_run.restype = C.c_double
        _lib.cpu_phys_run.argtypes = [C.POINTER(PhysT), C.POINTER(StateT), C.c_int, C.c_int, _dp, C.c_int]
        _lib.cpu_phys_run_many.restype = C.c_double
        _lib.cpu_pic_run.restype = C.c_double
        _lib.cpu_pic_run.argtypes = [C.POINTER(PhysT), C.POINTER(StateT), C.POINTER(BeamT), C.c_int, C.c_int, _dp,
                                     C.c_int]
        _lib.cpu_max_threads.restype = C.c_int
    return _lib


def _p(a, t=_dp):
    return a.ctypes.data_as(t)


def _ranges(mask, N):
    idx = np.nonzero(mask)[0]
    left = idx[idx < N // 2]
    right = idx[idx >= N // 2]
    lr = (int(left[0]), int(left[-1]) + 1) if len(left) else (0, 0)
    rr = (int(right[0]), int(right[-1]) + 1) if len(right) else (0, 0)
    return lr, rr


class CpuSim:
    """Holds the ctypes structs and numpy buffers for one model."""

    def __init__(self, m, J=None):
        N = m["N"]
        self.m = m
        self.keep = []
        pe = m["pmlE"].copy()
        pe[0] = pe[N] = False
        ph = m["pmlH"].copy()
        ph[N] = False
        (el0, el1), (er0, er1) = _ranges(pe, N)
        (hl0, hl1), (hr0, hr1) = _ranges(ph, N)
        arr = lambda a: self._keep(np.ascontiguousarray(a, dtype=np.float64))
        self.probes = np.ascontiguousarray(m["probes"], dtype=np.int32)
        self.probe_out = np.zeros((len(self.probes), m["steps"]))
        self.J = None if J is None else np.ascontiguousarray(J, dtype=np.float64)
        self.p = PhysT(N, m["steps"], m["Sc"], el0, el1, er0, er1, hl0, hl1, hr0, hr1,
                       _p(arr(m["bE"])), _p(arr(m["cE"])), _p(arr(m["bH"])), _p(arr(m["cH"])),
                       m["F"], m["R"], _p(arr(m["A"])), _p(arr(m["B"])), _p(arr(m["C"])),
                       _p(arr(m["eps"])), m["chi3"], m["src_cell"], _p(arr(m["src"])),
                       len(self.probes), _p(self.probes, _ip), _p(self.probe_out),
                       _p(self.J) if self.J is not None else _dp())
        ns = m["R"] - m["F"]
        self.E, self.H, self.psiE, self.psiH = (np.zeros(N + 1) for _ in range(4))
        self.D, self.P, self.Pm = (np.zeros(max(ns, 1)) for _ in range(3))
        self.s = StateT(*(_p(a) for a in (self.E, self.H, self.psiE, self.psiH, self.D, self.P, self.Pm)))
        self.scratch = np.zeros(N + 1)

    def _keep(self, a):
        self.keep.append(a)
        return a

    def run(self, t0=0, t1=None, threads=1):
        t1 = self.m["steps"] if t1 is None else t1
        return lib().cpu_phys_run(C.byref(self.p), C.byref(self.s), t0, t1, _p(self.scratch), threads)


def run_model(m, threads=1, steps=None):
    sim = CpuSim(m)
    secs = sim.run(0, m["steps"] if steps is None else steps, threads)
    return sim, secs


class CpuPic(CpuSim):
    """A model with a particle beam (oracle.pic.Beam's state, copied): the C
    restatement of oracle/pic.run.  One thread is bitwise oracle/pic.py;
    more threads push in parallel and sum the deposit in per-thread partials."""

    def __init__(self, m, beam, threads=1):
        super().__init__(m)
        N, n = m["N"], len(beam.x)
        self.threads = max(1, int(threads))
        self.x = np.ascontiguousarray(beam.x, dtype=np.float64).copy()
        self.u = np.ascontiguousarray(beam.u, dtype=np.float64).reshape(-1, 3).copy()
        self.alive = np.ascontiguousarray(beam.alive, dtype=np.int8).copy()
        self.Hprev = np.zeros(N + 1)
        self.Jb = np.zeros(N + 1)
        self.B1 = np.zeros(self.threads * (N + 1))
        self.B2 = np.zeros(self.threads * (N + 1))
        self.m0 = np.zeros(n, dtype=np.int32)
        self.w0, self.w1 = np.zeros(n), np.zeros(n)
        self.b = BeamT(n, _p(self.x), _p(self.u), self.alive.ctypes.data_as(C.POINTER(C.c_int8)), beam.a,
                       beam.jscale, _p(self.Hprev), _p(self.Jb), _p(self.B1), _p(self.B2), _p(self.m0, _ip),
                       _p(self.w0), _p(self.w1))

    def run(self, t0=0, t1=None, threads=None):
        t1 = self.m["steps"] if t1 is None else t1
        nth = self.threads if threads is None else min(int(threads), self.threads)
        return lib().cpu_pic_run(C.byref(self.p), C.byref(self.s), C.byref(self.b), t0, t1, _p(self.scratch), nth)
