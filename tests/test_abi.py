"""The C-ABI library loads and exports every function the public header
declares (no device calls: this runs without a GPU)."""
import ctypes
import os
import re
import subprocess

import numpy as np

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


def test_host_side_rejections(lulesh_gpu, oracle, tmp_path):
    """Bad configurations, meshes and files are refused on the host, before
    any device work (so this runs without a GPU)."""
    lib = lulesh_gpu.lib()
    BAD = lulesh_gpu.BAD_ARG
    h = ctypes.c_void_p()

    def cfg_with(**kw):
        c = lulesh_gpu.default_config()
        for k, v in kw.items():
            setattr(c, k, v)
        return c

    for kw in ({"nx": 0}, {"nx": 601}, {"ic_mode": 3}, {"num_reg": 0}, {"cost": -1},
               {"balance": 17}, {"max_cycles": -1}):
        assert lib.lulesh_create(ctypes.byref(cfg_with(**kw)), ctypes.byref(h)) == BAD, kw
    for ef in (0.0, -1.0, float("nan"), 1.0e7):
        rc = lib.lulesh_create_scaled(ctypes.byref(cfg_with()), ctypes.c_double(ef), 0,
                                      ctypes.byref(h))
        assert rc == BAD, ef

    # .lmesh files: missing, truncated, indices out of range, 2.0 ICs
    one = lulesh_gpu.default_config()
    one.ic_mode = 1
    good = ("1 8\n" + "".join(f"{x} {y} {z}\n" for z in (0, 1) for y in (0, 1)
                                for x in (0, 1)) +
            "0 1 3 2 4 5 7 6\n0 0 0 0 0 0\n4 0 2 4 6\n4 0 1 4 5\n4 0 1 2 3\n4 1 3 5 7\n")
    cases = {"missing": None, "truncated": good[:40],
             "node index": good.replace("0 1 3 2 4 5 7 6", "0 1 3 2 4 5 7 8"),
             "neighbour": good.replace("0 0 0 0 0 0", "0 0 0 0 0 1"),
             "set entry": good.replace("4 1 3 5 7", "4 1 3 5 9"),
             "counts": "0 8\n" + good[4:]}
    for name, text in cases.items():
        p = tmp_path / f"{name.replace(' ', '_')}.lmesh"
        if text is not None:
            p.write_text(text)
        rc = lib.lulesh_create_lmesh(str(p).encode(), ctypes.byref(one), ctypes.byref(h))
        assert rc == BAD, name
    m = oracle.parse_lmesh(good)  # the unbroken file is a valid one-element mesh
    assert m["nodelist"].shape == (1, 8) and len(m["free_surf"]) == 4
    p = tmp_path / "good.lmesh"
    p.write_text(good)
    two = lulesh_gpu.default_config()
    assert lib.lulesh_create_lmesh(str(p).encode(), ctypes.byref(two), ctypes.byref(h)) == BAD

    # ensembles: no members, mixed IC modes, no stop time, a negative cycle
    # cap are refused before any device work
    cfgs = (lulesh_gpu.Config * 2)(lulesh_gpu.default_config(), lulesh_gpu.default_config())
    b = ctypes.c_void_p()
    assert lib.lulesh_batch_create(cfgs, None, 0, ctypes.byref(b)) == BAD
    for field, value in (("ic_mode", 1), ("stoptime", 0.0), ("max_cycles", -1)):
        cfgs[1] = cfg_with(**{field: value})
        assert lib.lulesh_batch_create(cfgs, None, 2, ctypes.byref(b)) == BAD, field
    assert not b.value

    # the array form takes the file's count bounds (corner-major indices stay
    # in int32); the counts are refused before any array is read
    xs = np.zeros(8)
    ids = np.zeros(8, dtype=np.int32)
    xp = xs.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    ip = ids.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
    for ne, nn in (((1 << 26) + 1, 8), (1, (1 << 27) + 1), (0, 8), (1, 7)):
        rc = lib.lulesh_create_mesh(ne, nn, xp, xp, xp, ip, ip, 0, None, 0, None, 0, None,
                                    0, None, ctypes.byref(one), ctypes.byref(h))
        assert rc == BAD, (ne, nn)

    # the VTK writer's array form is for the n^3 cube
    z = np.zeros(8)
    dp = z.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    assert lib.lulesh_write_domain_arrays(str(tmp_path / "x.vtk").encode(), 7, 0,
                                          ctypes.c_double(0.0), *([dp] * 10)) == BAD
