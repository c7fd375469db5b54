"""The C-ABI library loads and exports every function the public header
declares (no device calls: this runs without a GPU)."""
import ctypes
import os
import re
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "lulesh_mi355x.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(lulesh_[a-z_0-9]+)\s*\(", text)))


def test_header_declares_the_abi():
    names = declared_functions()
    assert "lulesh_create" in names and "lulesh_step" in names
    assert len(names) >= 15


def test_library_exports_every_declared_symbol(lulesh_gpu):
    lib = lulesh_gpu.lib()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    # the Python binding covers the same set
    assert sorted(lulesh_gpu.EXPORTS) == declared_functions()


def test_exports_are_plain_c_symbols(lulesh_gpu):
    out = subprocess.run(["nm", "-D", "--defined-only", lulesh_gpu.LIB_PATH],
                         capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r" T (lulesh_[a-z_0-9]+)$", out, flags=re.M))
    assert set(declared_functions()) <= exported


def test_host_only_queries(lulesh_gpu):
    lib = lulesh_gpu.lib()
    assert lib.lulesh_kernel_count() >= 3
    names = [lib.lulesh_kernel_name(i).decode()
             for i in range(lib.lulesh_kernel_count())]
    assert all(n.startswith("k_") for n in names)
    cfg = lulesh_gpu.default_config()
    assert (cfg.nx, cfg.ic_mode, cfg.num_reg, cfg.balance, cfg.cost) == (30, 2, 11, 1, 1)
    assert cfg.stoptime == 1.0e-2
    # argument checks happen before any device work
    bad = lulesh_gpu.Config()
    ctypes.memmove(ctypes.byref(bad), ctypes.byref(cfg), ctypes.sizeof(cfg))
    bad.stoptime = -1.0
    h = ctypes.c_void_p()
    assert lib.lulesh_create(ctypes.byref(bad), ctypes.byref(h)) == lulesh_gpu.BAD_ARG
    assert lib.lulesh_step(None, 1, None) == lulesh_gpu.BAD_ARG
