"""The C-ABI library loads and exports every function include/*.h declares.

No compute calls are made (this runs without a GPU); the ctypes binding must
bind exactly the declared set.
"""
import ctypes
import glob
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        src = open(h).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\*?\s*(fdtd_[a-z0-9_]+|pic_[a-z0-9_]+|batch_[a-z0-9_]+|serve_[a-z0-9_]+)\s*\(",
                             src, flags=re.M):
            names.add(m.group(1))
    return names


def test_header_declares_functions():
    names = declared_functions()
    assert {"fdtd_ctx_create", "fdtd_add_sim", "fdtd_run", "fdtd_read_field", "fdtd_last_error"} <= names


def test_library_exports_every_declared_symbol():
    from fdtd_pic_amd import _capi
    lib = ctypes.CDLL(_capi.LIB_PATH)
    missing = [n for n in sorted(declared_functions()) if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_header_exactly():
    from fdtd_pic_amd import _capi
    assert set(_capi.SIGNATURES) == declared_functions()
    _capi.load()   # binds restype/argtypes for every symbol; no device call


def test_no_device_is_reported_cleanly():
    """Without a visible GPU the context constructor fails with an error code, not a crash."""
    import pytest
    from fdtd_pic_amd import _capi, engine
    try:
        import torch
        if torch.cuda.device_count() > 0:
            pytest.skip("a GPU is visible")
    except Exception:
        pass
    with pytest.raises(_capi.FdtdError):
        engine.Context("f64")
