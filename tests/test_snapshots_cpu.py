"""Snapshot archive (fdtd_pic_amd.snapshots): write -> read round trip of the
oracle's Ex_History (vidMake, Solver_Engine.py:57-68) with its geometry; the
archive loads without pickles; the VideoMaker drop-in writes it under
"Ex fields"."""
import numpy as np

from refcases import oracle_objects, small_setup


def _oracle_run():
    from oracle import faithful as FO
    setup = small_setup(Nz=300, T=400, F=200, R=299, nzsrc=60)
    V, P, C_V, C_P = oracle_objects(setup, 9e9, SineCont=True, TFSF=True)
    FO.controller(V, P, C_V, C_P)
    return V, P


def test_history_round_trip(tmp_path):
    from fdtd_pic_amd import snapshots
    V, P = _oracle_run()
    assert np.abs(V.Ex_History).max() > 0
    f = snapshots.write_history(str(tmp_path / "h.npz"), V, P)
    d = snapshots.read(f)
    assert np.array_equal(d["Ex_History"], V.Ex_History)
    assert d["interval"] == P.vidInterval and d["dt"] == P.delT and d["dz"] == P.dz
    assert tuple(d["slab"]) == (P.materialFrontEdge, P.materialRearEdge) and d["pml_width"] == P.pmlWidth
    np.testing.assert_array_equal(d["step"], np.arange(len(V.Ex_History)) * P.vidInterval)
    # a plain archive: numpy's loader with pickles refused reads every member
    with np.load(f, allow_pickle=False) as z:
        assert set(z.files) >= {"Ex_History", "step", "time_s", "z_m"}


def test_video_maker_writes_archive(tmp_path, monkeypatch):
    import Validation_Physics
    from fdtd_pic_amd import snapshots
    V, P = _oracle_run()
    monkeypatch.chdir(tmp_path)
    f = Validation_Physics.VideoMaker(P, V)
    assert f == str(tmp_path / "Ex fields" / "Ex_History.npz")
    assert np.array_equal(snapshots.read(f)["Ex_History"], V.Ex_History)


def test_rejects_foreign_archive(tmp_path):
    import pytest
    from fdtd_pic_amd import snapshots
    np.savez(tmp_path / "x.npz", format=np.array("other"), Ex_History=np.zeros((2, 3)))
    with pytest.raises(ValueError):
        snapshots.read(str(tmp_path / "x.npz"))
