"""The C-ABI library loads and exports every function declared in include/*.h
(CPU only: no compute call is made)."""
import ctypes
import glob
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = re.sub(r"/\*.*?\*/", "", open(h).read(), flags=re.S)
        names |= set(re.findall(r"\b(chk_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    import riggle_amd
    lib = ctypes.CDLL(riggle_amd.LIB_PATH)
    names = declared_functions()
    assert len(names) >= 10
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_argument_errors_are_reported_not_raised():
    import riggle_amd
    L = riggle_amd.lib()
    h = ctypes.c_void_p()
    rc = L.chk_index_create(0, 1, 0, 0, None, None, None, None, ctypes.byref(h))
    assert rc == 1 and b"nfiles" in L.chk_last_error()
    rc = L.chk_query_create(None, 1, 0, None, None, None, None, ctypes.byref(h))
    assert rc == 1
    assert L.chk_version().startswith(b"chuckle")
