"""The C-ABI library loads without a GPU and exports every function that
include/*.h declares."""
import ctypes as C
import os
import re

from conftest import ROOT


def _declared():
    names = set()
    for h in ("t3ns_amd.h", "t3ns_host.h"):
        with open(os.path.join(ROOT, "include", h)) as f:
            txt = f.read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        for m in re.finditer(r"\b(t3ns_[A-Za-z0-9_]+)\s*\(", txt):
            names.add(m.group(1))
    names.discard("t3ns_dev_matvec")
    return names


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(os.path.join(ROOT, "t3ns_amd", "libt3ns_amd.so"))
    names = _declared()
    assert len(names) >= 20
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing


def test_device_count_without_gpu_is_not_an_error():
    from t3ns_amd import lib
    assert lib().t3ns_device_count() >= 0
