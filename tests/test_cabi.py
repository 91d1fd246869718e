"""The C-ABI library loads and exports every symbol include/mldgpu.h declares (no device calls)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "mldgpu.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\s*\*)\s*(mld_[a-z0-9_]+)\s*\(", txt, re.M)))


def test_library_exports_every_declared_symbol():
    import mld_amd
    from mld_amd import build as b

    if not os.path.exists(mld_amd.LIBPATH):
        b.build(verbose=False)
    lib = ctypes.CDLL(mld_amd.LIBPATH)
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing
    assert set(declared_symbols()) == set(mld_amd.EXPORTS)


def test_no_cpu_fallback_without_library(monkeypatch, tmp_path):
    import mld_amd

    monkeypatch.setattr(mld_amd, "LIBPATH", str(tmp_path / "absent.so"))
    monkeypatch.setattr(mld_amd, "_lib", None)
    with pytest.raises(ImportError):
        mld_amd.load()
