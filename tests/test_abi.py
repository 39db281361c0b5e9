"""CPU-side checks of the C-ABI: the library loads without a GPU and exports
every function include/ibfs.h declares (no compute calls here)."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "ibfs.h")
LIB = os.path.join(ROOT, "ibfs-if2d_amd", "libibfs_hip.so")


def declared():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(ibfs_[a-z0-9_]+)\s*\(", txt)))


def test_header_declares_the_step_api():
    names = declared()
    for must in ("ibfs_create", "ibfs_advance", "ibfs_step", "ibfs_build_cholesky", "ibfs_set_cholesky",
                 "ibfs_set_state", "ibfs_get_state", "ibfs_destroy", "ibfs_last_error"):
        assert must in names


def test_library_exports_every_declared_symbol():
    assert os.path.exists(LIB), "build with __graft_entry__.build() first"
    lib = ctypes.CDLL(LIB)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_null_arguments_are_rejected_without_a_gpu():
    lib = ctypes.CDLL(LIB)
    lib.ibfs_destroy.argtypes = [ctypes.c_void_p]
    assert lib.ibfs_destroy(None) == 1
    lib.ibfs_dims.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    assert lib.ibfs_dims(None, None) == 1
