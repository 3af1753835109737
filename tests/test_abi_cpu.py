"""The C-ABI library loads on a machine without a GPU, exports every entry point
that include/uclbs_hip.h declares, and validates arguments before touching HIP."""
import ctypes
import os
import re

import pytest

from conftest import ROOT


def _declared():
    txt = open(os.path.join(ROOT, "include", "uclbs_hip.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(uclbs_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_header(pkg):
    lib = pkg.load_library()
    names = _declared()
    assert len(names) >= 15
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(pkg.EXPORTS)
    assert lib.uclbs_abi_version() == 1


def test_create_rejects_bad_config_without_gpu(pkg):
    lib = pkg.load_library()
    cfg = pkg.Config()
    cfg.lattice, cfg.precision, cfg.alpha_mode, cfg.nu = 19, 0, 0, 1e-3
    ctx = ctypes.c_void_p()
    assert lib.uclbs_create(ctypes.byref(cfg), ctypes.byref(ctx)) == -1
    assert b"lattice" in lib.uclbs_last_error()
    cfg.lattice, cfg.nu = 27, -1.0
    assert lib.uclbs_create(ctypes.byref(cfg), ctypes.byref(ctx)) == -1
    assert lib.uclbs_destroy(None) == 0
    assert lib.uclbs_sync(None) == -1


def test_no_cpu_fallback(pkg, tmp_path):
    with pytest.raises(pkg.UclbsError):
        pkg._lib = None
        try:
            pkg.load_library(str(tmp_path / "missing.so"))
        finally:
            pkg._lib = None
