"""The C ABI library loads and exports every symbol include/avbd.h declares."""
import ctypes
import re
from pathlib import Path

from avbd_amd import _native as N

ROOT = Path(__file__).resolve().parent.parent


def declared_symbols():
    text = (ROOT / "include" / "avbd.h").read_text()
    return sorted(set(re.findall(r"\b(avbd_[a-z_0-9]+)\s*\(", text)))


def test_header_declares_the_bound_symbols():
    decl = declared_symbols()
    assert len(decl) >= 20
    bound = {"avbd_" + k for k in N.SIGNATURES}
    assert set(decl) == bound, set(decl) ^ bound


def test_library_exports_every_declared_symbol():
    lib = N.load_library()
    for name in declared_symbols():
        assert hasattr(lib, name), f"{name} missing from {N.LIB_PATH}"


def test_library_links_only_hip_runtime():
    # The library must load on a host without a GPU and report zero devices
    # there instead of crashing.
    n = N.device_count()
    assert n >= 0


def test_struct_layout_matches_header():
    assert ctypes.sizeof(N.Params) == 4 * 17
    assert ctypes.sizeof(N.Report) == 4 * 13
    # 1 int + 11 ptrs + 1 int + 5 ptrs + 11 ptrs + 2 ptrs + 4 ints (with padding)
    ptr = ctypes.sizeof(ctypes.c_void_p)
    assert ctypes.sizeof(N.Scene) == 8 + 11 * ptr + 8 + 18 * ptr + 16
    # avbd_views: int + 4 ptrs + 2 ints + 3 ptrs + int + ptr + int (with padding)
    assert ctypes.sizeof(N.Views) == 8 + 4 * ptr + 8 + 3 * ptr + 8 + ptr + 8
    assert N.Views.lam.offset == 8 + 4 * ptr + 8 and N.Views.header.offset == 24 + 7 * ptr
