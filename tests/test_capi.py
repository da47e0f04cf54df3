"""The C ABI library loads and exports every symbol include/*.h declares; no
compute call is made (this runs without a GPU)."""
import ctypes as C
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    inc = os.path.join(ROOT, "include")
    for f in os.listdir(inc):
        if not f.endswith(".h"):
            continue
        txt = open(os.path.join(inc, f)).read()
        if "extern \"C\"" not in txt:
            continue
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[a-z_]+\s*\*?\s*(pfast_[a-z0-9_]+)\s*\(", txt, re.M):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol():
    import pfast_amd._lib as L
    lib = L.lib()
    names = declared_symbols()
    assert len(names) >= 20
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing
    # every declared entry point is bound with a signature in the Python mirror
    assert names <= set(L.exported_symbols()), sorted(names - set(L.exported_symbols()))


def test_version_and_device_count_without_gpu():
    import pfast_amd._lib as L
    lib = L.lib()
    assert lib.pfast_version() >= 10000
    n = C.c_int(-1)
    assert lib.pfast_device_count(C.byref(n)) == 0
    assert n.value >= 0
