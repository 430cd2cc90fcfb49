"""Drop-in for the reference's Validation_Physics.py (VideoMaker).

The reference turns V.Ex_History into one PNG per row in an "Ex fields"
folder, then an MP4 (Validation_Physics.py:18-71; it deletes that folder
first).  Here VideoMaker writes the same rows, with the geometry its frames
draw (slab, PML, probe cells, the step of each row), to
"Ex fields/Ex_History.npz" (fdtd_pic_amd.snapshots) and returns the path:
the frames' data, not images.  Rendering (matplotlib, moviepy, natsort) is
outside this package.
"""
import os

from fdtd_pic_amd import snapshots


def VideoMaker(P, V, folder="Ex fields"):
    path = os.path.join(os.getcwd(), folder)
    os.makedirs(path, exist_ok=True)
    return snapshots.write_history(os.path.join(path, "Ex_History.npz"), V, P)
