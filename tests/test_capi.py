"""CPU: libgvx.so loads and exports every entry point include/gvx.h declares (no GPU calls)."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(gvx_\w+)\s*\(", text, flags=re.M)))


def test_header_declares_entry_points():
    names = declared("gvx.h")
    assert len(names) >= 40
    assert "gvx_dist_pba" in names and "gvx_batch_collide" in names


def test_library_exports_all_declared_symbols():
    import gvx
    assert os.path.exists(gvx.LIB_PATH), "libgvx.so missing: run gpu-voxels_amd/build.py"
    lib = ctypes.CDLL(gvx.LIB_PATH)
    missing = [n for n in declared("gvx.h") if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_every_symbol():
    import gvx
    src = open(gvx.__file__).read()
    missing = [n for n in declared("gvx.h") if '"%s"' % n not in src]
    assert not missing, missing
