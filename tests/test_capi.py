"""The C-ABI library loads and exports every entry point include/mfdfd.h
declares, with a binding for each (no device compute is called here)."""
import ctypes as C
import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
HEADER = os.path.join(os.path.dirname(HERE), "include", "mfdfd.h")


def declared():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(mfd_[A-Za-z0-9_]+)\s*\(", txt)))


def test_header_declares_entry_points():
    names = declared()
    assert "mfd_curl_csc" in names and "mfd_solve" in names and len(names) >= 25


def test_library_exports_every_declared_symbol(pkg):
    L = pkg.lib()
    raw = C.CDLL(pkg.LIB_PATH)
    for n in declared():
        assert hasattr(raw, n), f"{n} not exported"
        assert n in pkg.SIGNATURES, f"{n} has no binding"
    assert L.mfd_version().startswith(b"libmfdfd")


def test_errors_are_reported_without_gpu(pkg):
    import pytest
    with pytest.raises(pkg.MfdArgumentError):
        pkg.distweights(5.0, 0, (0, 1), [0.0, 0.5], [0.5, 0.5], False)
