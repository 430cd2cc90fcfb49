"""On-disk snapshot output: Ex_History rows and the geometry needed to plot them.

The reference fills V.Ex_History with one row of Ex every P.vidInterval steps
(vidMake, Solver_Engine.py:57-68; rows allocated at MasterController.py:160)
and renders the rows to PNG frames and an MP4 (Validation_Physics.VideoMaker,
Validation_Physics.py:18-71).  On the device the rows are written by the step
kernel itself (snapshot rows of the simulation block, read back with
fdtd_read_snapshots).  This module stores them, with what VideoMaker draws
around them (slab edges, PML width, probe cells, time of each row), as an
uncompressed .npz archive: plain arrays, loadable with numpy.load and no
pickling.  Rendering to images or video is left to the caller.
"""
import numpy as np

FORMAT = "fdtd-snapshots-1"


def write(path, rows, *, interval, dz, dt, slab=(0, 0), pml_width=0, probes=(), first_row_step=0):
    """Write snapshot rows (rows x (N+1), Ex along the grid) and their geometry.

    Row r holds Ex after step first_row_step + r * interval (the reference's
    row r is step r * vidInterval; its row 0 is never written and stays 0)."""
    rows = np.ascontiguousarray(rows, dtype=np.float64)
    if rows.ndim != 2:
        raise ValueError("rows must be a 2-D array (rows x cells)")
    steps = first_row_step + np.arange(rows.shape[0], dtype=np.int64) * int(interval)
    np.savez(path, format=np.array(FORMAT), Ex_History=rows, step=steps, time_s=steps * float(dt),
             z_m=np.arange(rows.shape[1], dtype=np.float64) * float(dz), interval=np.int64(interval),
             dz=np.float64(dz), dt=np.float64(dt), slab=np.asarray(slab, dtype=np.int64),
             pml_width=np.int64(pml_width), probes=np.asarray(probes, dtype=np.int64))
    return path


def read(path):
    """The arrays written by write(), as a dict (numpy.load without pickles)."""
    with np.load(path, allow_pickle=False) as z:
        if str(z["format"]) != FORMAT:
            raise ValueError(f"{path}: not a {FORMAT} archive")
        return {k: z[k] for k in z.files if k != "format"}


def write_history(path, V, P):
    """V.Ex_History of a reference-shaped run (the drop-in integrators fill it
    from the device rows) with the reference's geometry (P.vidInterval,
    P.dz, P.delT, the slab, PML and probe cells VideoMaker draws)."""
    return write(path, V.Ex_History, interval=P.vidInterval, dz=P.dz, dt=P.delT,
                 slab=(P.materialFrontEdge, P.materialRearEdge), pml_width=P.pmlWidth,
                 probes=(P.x1Loc, P.x2Loc))


def write_sim(path, sim, dz, dt):
    """The snapshot rows of a device simulation (engine.Sim created with
    snap_interval / snap_rows set) and its geometry."""
    s = sim.spec
    if not s.snap_rows:
        raise ValueError("the simulation records no snapshots (snap_rows = 0)")
    return write(path, sim.snapshots(), interval=s.snap_interval, dz=dz, dt=dt, slab=(s.F, s.R),
                 pml_width=s.Ls, probes=[c for c, _, _ in s.probes])
