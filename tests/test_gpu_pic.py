"""PIC coupling on the GPU against the FP64 oracle (oracle/pic.py).

Cases (config-1 slab, thermal electron beam drifting at 0.1c):
  beam_only  no source: every field is radiated by the beam current
  driven     a 1e9 V/m pulse crosses the beam before reaching the slab
  absorbed   a beam drifting into the left boundary (particles parked)
  absorbed_right  the same at the right boundary
FP64: fields, probes, particle displacement and u within 1e-10 relative L2;
FP32: within 1e-5; the same for one-cell groups at mixed cell densities
(push/count's register and LDS-held chunks).  Also: bitwise invariance of the run under a permutation
of the input particles, agreement between sort/deposit wave widths, the
fused field + scatter launch bitwise equal to separate launches, and a
reduced config 3 (N = 65,536, 1e5 particles, 300 steps) in FP32.
"""
import numpy as np
import pytest

from refcases import rel_l2

pytestmark = pytest.mark.gpu


def _params(kind, N=4096, steps=600, n=20000):
    from fdtd_pic_amd import configs
    p = configs.base(N, steps, slab_start=N // 2, slab_stop=N // 2 + 400)
    C0 = configs.C0
    lo, hi, vd = 1200, 1800, 0.1 * C0
    if kind == "beam_only":
        p["src_amp"] = 0.0
    elif kind == "driven":
        p["src_amp"] = 1e9
    elif kind == "absorbed":
        p["src_amp"] = 0.0
        p["steps"] = 250               # about half of the beam crosses cell 2
        lo, hi, vd = 3, 60, -0.3 * C0
    elif kind == "absorbed_right":
        p["src_amp"] = 0.0
        p["steps"] = 250               # about half of the beam crosses cell N-3
        lo, hi, vd = N - 60, N - 3, 0.3 * C0
    p["beam"] = dict(seed=17, n=n, cell_lo=lo, cell_hi=hi, v_drift=vd, v_th=1e-3 * C0, density=1e16)
    return p


def _oracle(p):
    from fdtd_pic_amd import configs
    from oracle import pic as P, physical as ph
    m = ph.build(p)
    x, u, q, mm, w = configs.beam_inputs(p)
    b = P.Beam(x, u, q, mm, w, m["dt"], p["N"], p["Sc"])
    s, probes = P.run(m, b)
    return dict(E=s.E, H=s.H, probes=probes, x=b.x, u=b.u, alive=b.alive, x0=x)


def _oracle_inputs(p):
    from fdtd_pic_amd import configs
    return configs.beam_inputs(p)[0]


def _device(p, prec, group=0, perm=None):
    from fdtd_pic_amd import configs, engine
    x, u, q, mm, w = configs.beam_inputs(p)
    if perm is not None:
        x, u = x[perm], u[perm]
    dt = p["Sc"] * p["dz"] / configs.C0
    with engine.Context(prec) as ctx:
        sim = ctx.add(engine.physical_spec(p), beam=engine.Beam(x, u, q, mm, w, dt, p["dz"], group))
        ctx.run_all()
        xo, uo, alive = sim.particles()
        return dict(E=sim.read("E"), H=sim.read("H"), probes=sim.probes(), x=xo, u=uo, alive=alive,
                    J=sim.read("J"))


@pytest.fixture(scope="module")
def oracles():
    return {k: (_params(k), _oracle(_params(k))) for k in ("beam_only", "driven", "absorbed", "absorbed_right")}


@pytest.mark.parametrize("kind", ["beam_only", "driven", "absorbed", "absorbed_right"])
@pytest.mark.parametrize("prec,tol", [("f64", 1e-10), ("f32", 1e-5)])
def test_pic_parity(oracles, kind, prec, tol):
    p, o = oracles[kind]
    print(f"pic parity {kind} {prec}", flush=True)
    d = _device(p, prec)
    errs = dict(E=rel_l2(d["E"], o["E"]), H=rel_l2(d["H"], o["H"]), probes=rel_l2(d["probes"], o["probes"]),
                dx=rel_l2(d["x"] - o["x0"], o["x"] - o["x0"]), u=rel_l2(d["u"], o["u"]))
    print(kind, prec, {k: f"{v:.2e}" for k, v in errs.items()}, "absorbed", int((~o["alive"]).sum()), flush=True)
    assert np.abs(o["E"]).max() > 0
    assert np.array_equal(d["alive"], o["alive"])
    if kind.startswith("absorbed"):
        assert 100 < (~o["alive"]).sum() < len(o["alive"]) - 100
    for k, v in errs.items():
        assert v <= tol, (kind, prec, k, v)


@pytest.mark.parametrize("prec,tol", [("f64", 1e-10), ("f32", 1e-5)])
def test_one_cell_groups_mixed_density(prec, tol):
    """One-cell groups (push/count writes the new state) on a beam of about
    125 particles per cell: cells of up to 128 particles are held in
    registers, fuller ones (up to 158 here) push a third chunk whose state
    waits in LDS, in one run; against the oracle at the parity tolerances.
    Cells above 192 (the counting pass plus the writing pass) are the
    absorbed cases of test_pic_parity (about 350 per cell)."""
    p = _params("driven", steps=300, n=25000)
    p["beam"].update(cell_lo=1200, cell_hi=1400)
    o = _oracle(p)
    x0 = o["x0"]
    cells = np.bincount(np.floor(x0).astype(int))
    assert cells.max() > 128 and (cells[1200:1400] <= 128).any()
    print(f"one-cell groups {prec}", flush=True)
    d = _device(p, prec, group=1)
    errs = dict(E=rel_l2(d["E"], o["E"]), H=rel_l2(d["H"], o["H"]), probes=rel_l2(d["probes"], o["probes"]),
                dx=rel_l2(d["x"] - x0, o["x"] - x0), u=rel_l2(d["u"], o["u"]))
    print("one-cell groups", prec, {k: f"{v:.2e}" for k, v in errs.items()}, flush=True)
    assert np.array_equal(d["alive"], o["alive"])
    for k, v in errs.items():
        assert v <= tol, (prec, k, v)


@pytest.mark.parametrize("prec,tol", [("f64", 1e-10), ("f32", 1e-5)])
def test_one_cell_lds_rows_bitwise(prec, tol):
    """LDS rows per cell of the one-cell kernel (fdtd_set_pic_lds_rows) only
    change where a chunk's new state waits: a beam of about 350 particles per
    cell (the absorbed case) pushed with one row (cells above 192 take the
    counting pass and the writing pass), with the rows chosen from its density
    (one pass where they suffice) and with three rows gives bitwise the same
    fields, J and particles, and matches the oracle."""
    from fdtd_pic_amd import configs, engine
    p = _params("absorbed")
    x, u, q, mm, w = configs.beam_inputs(p)
    assert np.bincount(np.floor(x).astype(int)).max() > 320
    dt = p["Sc"] * p["dz"] / configs.C0
    outs = []
    for rows in (1, 0, 3):
        print("one-cell rows", prec, rows, flush=True)
        with engine.Context(prec) as ctx:
            ctx.set_pic_lds_rows(rows)
            sim = ctx.add(engine.physical_spec(p), beam=engine.Beam(x, u, q, mm, w, dt, p["dz"], 1))
            ctx.run_all()
            xo, uo, alive = sim.particles()
            outs.append(dict(E=sim.read("E"), H=sim.read("H"), J=sim.read("J"), probes=sim.probes(), x=xo, u=uo,
                             alive=alive))
    for o in outs[1:]:
        for k in outs[0]:
            assert np.array_equal(o[k], outs[0][k]), k
    ref = _oracle(p)
    assert rel_l2(outs[0]["E"], ref["E"]) <= tol and rel_l2(outs[0]["x"] - x, ref["x"] - x) <= tol


def test_permutation_invariance_bitwise():
    p = _params("driven", steps=200)
    rng = np.random.default_rng(5)
    perm = rng.permutation(p["beam"]["n"])
    a = _device(p, "f32")
    b = _device(p, "f32", perm=perm)
    for k in ("E", "H", "J", "probes"):
        assert np.array_equal(a[k], b[k]), k
    assert np.array_equal(b["x"], a["x"][perm]) and np.array_equal(b["u"], a["u"][perm])


def test_wave_width_independence():
    """Cells per push/count group change only the order of the FP64 CIC sums.
    Groups of 3 cells (~100 particles) take push/count's one-pass path, 5 and
    64 its two passes, 1 the one-cell kernel."""
    p = _params("driven", steps=300)
    ref = _device(p, "f64", group=0)
    x0 = _oracle_inputs(p)
    for g in (1, 3, 5, 64):
        d = _device(p, "f64", group=g)
        assert rel_l2(d["E"], ref["E"]) < 1e-13, g
        assert rel_l2(d["x"] - x0, ref["x"] - x0) < 1e-13, g


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("K", [1, 2])
def test_fused_scatter_bitwise(prec, K):
    """The scatter run in the field step's launch gives bitwise the same run as
    separate launches.  Two beams in one context, one with a Lorentz slab and
    one in vacuum (two tile groups: one field launch alone, one fused), the
    second absorbed at the right boundary and finishing 50 steps later (the
    work lists are rebuilt mid-run)."""
    from fdtd_pic_amd import configs, engine
    pa = _params("driven", steps=200)
    pb = _params("absorbed_right")
    pb["slab_stop"] = pb["slab_start"]

    def run(fuse):
        print(f"fused scatter {prec} K={K} fuse={fuse}", flush=True)
        out = []
        with engine.Context(prec, K=K) as ctx:
            ctx.set_pic_fusion(fuse)
            sims = []
            for p in (pa, pb):
                x, u, q, mm, w = configs.beam_inputs(p)
                dt = p["Sc"] * p["dz"] / configs.C0
                sims.append(ctx.add(engine.physical_spec(p), beam=engine.Beam(x, u, q, mm, w, dt, p["dz"])))
            ctx.run_all()
            for s in sims:
                xo, uo, alive = s.particles()
                out.append(dict(E=s.read("E"), H=s.read("H"), J=s.read("J"), probes=s.probes(), x=xo, u=uo,
                                alive=alive))
        return out

    a, b = run(True), run(False)
    assert (~b[1]["alive"]).sum() > 100
    for i in range(2):
        for k in a[i]:
            assert np.array_equal(a[i][k], b[i][k]), (i, k)


def test_reduced_config3_f32():
    from fdtd_pic_amd import configs
    p = configs.config3(steps=300, n_particles=100_000)
    print("reduced config3 oracle", flush=True)
    o = _oracle(p)
    d = _device(p, "f32")
    errs = dict(E=rel_l2(d["E"], o["E"]), H=rel_l2(d["H"], o["H"]),
                dx=rel_l2(d["x"] - o["x0"], o["x"] - o["x0"]), u=rel_l2(d["u"], o["u"]))
    print("config3 reduced f32", {k: f"{v:.2e}" for k, v in errs.items()}, flush=True)
    for k, v in errs.items():
        assert v <= 1e-5, (k, v)


def test_config3_full_size_invariants():
    """Config 3 at full size (1e6 electrons, 20,000 steps, FP64), checked by
    properties that need no oracle run: no particle is lost (the beam stays
    clear of the boundaries), the deposited current sums to the particles'
    q w beta_x (CIC weights of a particle sum to one), and a second run is
    bitwise identical."""
    from fdtd_pic_amd import configs

    def run():
        p = configs.config3()
        d = _device(p, "f64")
        return p, d

    p, a = run()
    x, u, alive = a["x"], a["u"], a["alive"]
    N = p["N"]
    assert alive.all() and np.all((x >= 2) & (x < N - 2))
    # J' = Z0 q w c * sum of CIC weights of beta_x; the last deposit used the final u
    _, _, q, _, w = configs.beam_inputs(p)
    jscale = 376.730313668 * q * w * configs.C0         # Z0 q w c, as the library forms it
    bx = u[:, 0] / np.sqrt(1.0 + np.sum(u * u, axis=1))
    total, ref, mag = a["J"].sum(), jscale * bx.sum(), abs(jscale) * np.abs(bx).sum()
    print("sum J'", total, "q w c Z0 sum beta_x", ref, "scale", mag, flush=True)
    assert abs(total - ref) <= 1e-12 * mag
    _, b = run()
    for k in ("E", "H", "J", "x", "u"):
        assert np.array_equal(a[k], b[k]), k

