"""The C-ABI library loads and exports every entry point include/bes_hip.h declares.

No GPU call is made: only symbol lookup (dlsym) on the built library.
"""
import ctypes
import os

import pytest

from bes_amd import core


def test_header_declares_entry_points():
    names = core.exported_symbols()
    for must in ("bes_ctx_create", "bes_ctx_create_msh2", "bes_set_state", "bes_compute_residual",
                 "bes_explicit_run", "bes_flux_batch", "bes_jacobian_batch"):
        assert must in names


def test_library_exports_all_declared_symbols():
    if not os.path.exists(core.LIB_PATH):
        pytest.fail("libbes_hip.so not built: run python -m bes_amd.build")
    lib = ctypes.CDLL(core.LIB_PATH)
    missing = [n for n in core.exported_symbols() if not hasattr(lib, n)]
    assert not missing, missing


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    monkeypatch.setattr(core, "LIB_PATH", str(tmp_path / "nope.so"))
    monkeypatch.setattr(core, "_lib", None)
    with pytest.raises(RuntimeError, match="missing"):
        core.lib()
