"""TEST INFRASTRUCTURE: ctypes binding of the C oracle and readers for the
reference dumps.

The oracle (oracle/lulesh_oracle.c) restates the LULESH Lagrange step of
reference/C/LULESH-OMP.cc (mode 1) and reference/C/lulesh-2.0.3 (mode 2).
Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use
this module; the product never imports it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from dataclasses import dataclass, field

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liblulesh_oracle.so")
REF_DIR = os.path.join(HERE, "_ref")

NODE_FIELDS = ("x", "y", "z", "xd", "yd", "zd", "nodalMass")
ELEM_FIELDS = ("e", "p", "q", "ql", "qq", "v", "volo", "delv", "vdov",
               "arealg", "ss", "elemMass")

_lib = None


def build(quiet: bool = True) -> None:
    """Builds the oracle library (and oracle/_ref when the reference tree is
    present) with oracle/Makefile."""
    out = subprocess.run(["make", "-C", HERE, "all"], capture_output=True,
                         text=True)
    if out.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + out.stdout + out.stderr)
    if not quiet:
        print(out.stdout)


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.c_void_p
        D = ctypes.POINTER(ctypes.c_double)
        I = ctypes.POINTER(ctypes.c_int)
        L.lo_create.restype = P
        L.lo_create.argtypes = [ctypes.c_int] * 5
        L.lo_create_scaled.restype = P
        L.lo_create_scaled.argtypes = [ctypes.c_int] * 5 + [ctypes.c_double]
        IPT = ctypes.POINTER(ctypes.c_int)
        L.lo_create_mesh.restype = P
        L.lo_create_mesh.argtypes = ([ctypes.c_int, ctypes.c_int, D, D, D, IPT, IPT] +
                                     [ctypes.c_int, IPT] * 4)
        L.lo_destroy.argtypes = [P]
        L.lo_set_cbrt_mode.argtypes = [P, ctypes.c_int]
        L.lo_set_stoptime.argtypes = [P, ctypes.c_double]
        L.lo_set_threads.argtypes = [ctypes.c_int]
        L.lo_run.argtypes = [P, ctypes.c_int]
        L.lo_run.restype = ctypes.c_int
        L.lo_time_increment.argtypes = [P]
        L.lo_lagrange_nodal.argtypes = [P]
        L.lo_lagrange_nodal.restype = ctypes.c_int
        L.lo_lagrange_elements.argtypes = [P]
        L.lo_lagrange_elements.restype = ctypes.c_int
        L.lo_time_constraints.argtypes = [P]
        L.lo_num_elem.argtypes = [P]
        L.lo_num_node.argtypes = [P]
        L.lo_get.argtypes = [P, ctypes.c_char_p, D]
        L.lo_get.restype = ctypes.c_int
        L.lo_set.argtypes = [P, ctypes.c_char_p, D]
        L.lo_set.restype = ctypes.c_int
        L.lo_get_int.argtypes = [P, ctypes.c_char_p, I]
        L.lo_get_int.restype = ctypes.c_int
        L.lo_scalar.argtypes = [P, ctypes.c_char_p]
        L.lo_scalar.restype = ctypes.c_double
        L.lo_set_scalar.argtypes = [P, ctypes.c_char_p, ctypes.c_double]
        L.lo_elem_volume.argtypes = [D, D, D]
        L.lo_elem_volume.restype = ctypes.c_double
        L.lo_elem_volume_derivative.argtypes = [D] * 6
        L.lo_elem_node_normals.argtypes = [D] * 6
        L.lo_elem_shape_derivatives.argtypes = [D] * 4
        L.lo_elem_shape_derivatives.restype = ctypes.c_double
        L.lo_elem_hourglass_force.argtypes = [D] * 6 + [ctypes.c_double] * 2 + [D] * 3
        L.lo_cbrt_project.argtypes = [ctypes.c_double]
        L.lo_cbrt_project.restype = ctypes.c_double
        L.lo_rand_sequence.argtypes = [ctypes.c_uint, ctypes.c_int, I]
        _lib = L
    return _lib


def dptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def iptr(a: np.ndarray):
    assert a.dtype == np.int32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


def parse_lmesh(text: str) -> dict:
    """The lists of a .lmesh file (the format read by reference/CUDA/
    lulesh-kepler-singlegpu/lulesh.cu:728-848)."""
    tok = text.split()
    pos = 0

    def take(n, conv):
        nonlocal pos
        out = [conv(t) for t in tok[pos:pos + n]]
        if len(out) != n:
            raise ValueError("truncated .lmesh")
        pos += n
        return out

    ne, nn = take(2, int)
    xyz = np.array(take(3 * nn, float)).reshape(nn, 3)
    mesh = {"x": xyz[:, 0].copy(), "y": xyz[:, 1].copy(), "z": xyz[:, 2].copy(),
            "nodelist": np.array(take(8 * ne, int), np.int32).reshape(ne, 8),
            "neighbors": np.array(take(6 * ne, int), np.int32).reshape(ne, 6)}
    for name in ("symm_x", "symm_y", "symm_z", "free_surf"):
        (n,) = take(1, int)
        mesh[name] = np.array(take(n, int), np.int32)
    return mesh


def cube_mesh(nx: int) -> dict:
    """The Sedov cube nx^3 written as explicit lists (the structured
    connectivity and node sets of reference/C/LULESH-OMP.cc:2996-3151); the
    far faces' nodes form the free-surface set."""
    en = nx + 1
    k, j, i = np.meshgrid(np.arange(en), np.arange(en), np.arange(en), indexing="ij")
    coord = np.array([0.0] + [1.125 * c / nx for c in range(1, en)])
    x, y, z = coord[i].ravel(), coord[j].ravel(), coord[k].ravel()
    ek, ej, ei = [a.ravel() for a in np.meshgrid(np.arange(nx), np.arange(nx),
                                                  np.arange(nx), indexing="ij")]
    n0 = (ek * en + ej) * en + ei
    pl = en * en
    nodelist = np.stack([n0, n0 + 1, n0 + en + 1, n0 + en, n0 + pl, n0 + pl + 1,
                         n0 + pl + en + 1, n0 + pl + en], axis=1).astype(np.int32)
    el = np.arange(nx ** 3)
    nb = np.stack([np.where(ei > 0, el - 1, el), np.where(ei < nx - 1, el + 1, el),
                   np.where(ej > 0, el - nx, el), np.where(ej < nx - 1, el + nx, el),
                   np.where(ek > 0, el - nx * nx, el),
                   np.where(ek < nx - 1, el + nx * nx, el)], axis=1).astype(np.int32)
    nid = np.arange(en ** 3)
    far = (i.ravel() == nx) | (j.ravel() == nx) | (k.ravel() == nx)
    return {"x": x, "y": y, "z": z, "nodelist": nodelist, "neighbors": nb,
            "symm_x": nid[i.ravel() == 0].astype(np.int32),
            "symm_y": nid[j.ravel() == 0].astype(np.int32),
            "symm_z": nid[k.ravel() == 0].astype(np.int32),
            "free_surf": nid[far].astype(np.int32)}


class Oracle:
    """One oracle domain (a Sedov cube of nx^3 elements, or an unstructured
    mesh through from_mesh)."""

    @classmethod
    def from_mesh(cls, mesh: dict, cbrt_mode: int = 0, stoptime: float | None = None,
                  threads: int = 0) -> "Oracle":
        L = lib()
        L.lo_set_threads(threads)
        a = {k: np.ascontiguousarray(mesh[k], np.float64) for k in ("x", "y", "z")}
        ints = {k: np.ascontiguousarray(mesh[k], np.int32)
                for k in ("nodelist", "neighbors", "symm_x", "symm_y", "symm_z",
                          "free_surf")}
        ne = len(ints["nodelist"].reshape(-1, 8))
        h = L.lo_create_mesh(ne, len(a["x"]), dptr(a["x"]), dptr(a["y"]), dptr(a["z"]),
                             iptr(ints["nodelist"]), iptr(ints["neighbors"]),
                             len(ints["symm_x"]), iptr(ints["symm_x"]),
                             len(ints["symm_y"]), iptr(ints["symm_y"]),
                             len(ints["symm_z"]), iptr(ints["symm_z"]),
                             len(ints["free_surf"]), iptr(ints["free_surf"]))
        if not h:
            raise ValueError("lo_create_mesh failed")
        o = cls.__new__(cls)
        o.h = h
        o.nx, o.mode = 0, 1
        L.lo_set_cbrt_mode(h, cbrt_mode)
        if stoptime is not None:
            L.lo_set_stoptime(h, stoptime)
        o.num_elem = L.lo_num_elem(h)
        o.num_node = L.lo_num_node(h)
        return o

    def __init__(self, nx: int, mode: int = 2, num_reg: int = 11,
                 balance: int = 1, cost: int = 1, cbrt_mode: int = 0,
                 stoptime: float | None = None, threads: int = 0,
                 efactor: float = 1.0):
        L = lib()
        L.lo_set_threads(threads)
        self.h = L.lo_create_scaled(nx, mode, num_reg, balance, cost, efactor)
        if not self.h:
            raise ValueError("lo_create failed")
        self.nx, self.mode = nx, mode
        L.lo_set_cbrt_mode(self.h, cbrt_mode)
        if stoptime is not None:
            L.lo_set_stoptime(self.h, stoptime)
        self.num_elem = L.lo_num_elem(self.h)
        self.num_node = L.lo_num_node(self.h)

    def close(self):
        if self.h:
            lib().lo_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def run(self, max_cycles: int = 9999999) -> int:
        return lib().lo_run(self.h, max_cycles)

    def get(self, name: str) -> np.ndarray:
        n = self.num_node if name in ("x", "y", "z", "xd", "yd", "zd", "xdd",
                                      "ydd", "zdd", "fx", "fy", "fz",
                                      "nodalMass") else self.num_elem
        out = np.empty(n, np.float64)
        if lib().lo_get(self.h, name.encode(), dptr(out)) != n:
            raise KeyError(name)
        return out

    def set(self, name: str, values: np.ndarray) -> None:
        v = np.ascontiguousarray(values, np.float64)
        if lib().lo_set(self.h, name.encode(), dptr(v)) < 0:
            raise KeyError(name)

    def get_int(self, name: str, n: int | None = None) -> np.ndarray:
        if n is None:
            n = self.num_elem * (8 if name == "nodelist" else 1)
        out = np.empty(n, np.int32)
        if lib().lo_get_int(self.h, name.encode(), iptr(out)) < 0:
            raise KeyError(name)
        return out

    def scalar(self, name: str) -> float:
        return lib().lo_scalar(self.h, name.encode())

    def set_scalar(self, name: str, value: float) -> None:
        lib().lo_set_scalar(self.h, name.encode(), float(value))

    @property
    def cycle(self) -> int:
        return int(self.scalar("cycle"))

    def state(self) -> "State":
        s = State(nx=self.nx, cycle=self.cycle, time=self.scalar("time"),
                  deltatime=self.scalar("deltatime"),
                  dtcourant=self.scalar("dtcourant"),
                  dthydro=self.scalar("dthydro"))
        for f in NODE_FIELDS + ELEM_FIELDS:
            s.fields[f] = self.get(f)
        return s


@dataclass
class State:
    nx: int
    cycle: int
    time: float
    deltatime: float
    dtcourant: float
    dthydro: float
    fields: dict = field(default_factory=dict)
    reg_num_list: np.ndarray | None = None


def read_dump(path: str) -> State:
    """Reads a dump written by oracle/ref_dump.h."""
    raw = open(path, "rb").read()
    if raw[:8] != b"LULDUMP1":
        raise ValueError("not a LULESH dump: " + path)
    nx, ne, nn, cycle = np.frombuffer(raw, np.int32, 4, 8)
    sc = np.frombuffer(raw, np.float64, 4, 24)
    s = State(nx=int(nx), cycle=int(cycle), time=float(sc[0]),
              deltatime=float(sc[1]), dtcourant=float(sc[2]),
              dthydro=float(sc[3]))
    off = 56
    for f in NODE_FIELDS:
        s.fields[f] = np.frombuffer(raw, np.float64, int(nn), off).copy()
        off += 8 * int(nn)
    for f in ELEM_FIELDS:
        s.fields[f] = np.frombuffer(raw, np.float64, int(ne), off).copy()
        off += 8 * int(ne)
    s.reg_num_list = np.frombuffer(raw, np.int32, int(ne), off).copy()
    return s


def run_reference(mode: int, nx: int, its: int | None = None,
                  stoptime: float | None = None, threads: int = 4,
                  dump_path: str | None = None, extra_args=(),
                  timeout: float = 600) -> tuple[str, State | None]:
    """Runs a binary built from the reference sources (oracle/_ref) and
    returns its stdout and, when dump_path is given, its final state."""
    env = dict(os.environ)
    env["OMP_NUM_THREADS"] = str(max(2, threads))
    if dump_path:
        env["LULESH_REF_DUMP"] = dump_path
    if mode == 1:
        exe = os.path.join(REF_DIR, "lulesh_omp_dump" if dump_path else "lulesh_omp")
        args = [exe, str(nx)]
        if stoptime is not None:
            env["LULESH_REF_STOPTIME"] = repr(float(stoptime))
        if its is not None:
            raise ValueError("LULESH 1.0 has no iteration limit")
    else:
        exe = os.path.join(REF_DIR, "lulesh2_dump" if dump_path else "lulesh2")
        args = [exe, "-s", str(nx)]
        if its is not None:
            args += ["-i", str(its)]
        if stoptime is not None:
            raise ValueError("LULESH 2.0 run length is set with -i")
    args += list(extra_args)
    out = subprocess.run(args, capture_output=True, text=True, env=env,
                         timeout=timeout)
    if out.returncode != 0:
        raise RuntimeError(f"{args} failed ({out.returncode}):\n{out.stderr}")
    st = read_dump(dump_path) if dump_path else None
    return out.stdout, st


def parse_reference_output(text: str) -> dict:
    """Parses the verification block both reference programs print
    (reference/C/LULESH-OMP.cc:3195-3223, lulesh-2.0.3/lulesh-util.cc:
    180-216)."""
    res = {}
    for line in text.splitlines():
        s = line.strip()
        if s.startswith("Iteration count"):
            res["cycles"] = int(s.split("=")[1])
        elif s.startswith("Final Origin Energy"):
            res["origin_energy"] = float(s.split("=")[1])
        elif s.startswith("MaxAbsDiff"):
            res["max_abs_diff"] = float(s.split("=")[1])
        elif s.startswith("TotalAbsDiff"):
            res["total_abs_diff"] = float(s.split("=")[1])
        elif s.startswith("MaxRelDiff"):
            res["max_rel_diff"] = float(s.split("=")[1])
        elif s.startswith("Elapsed time"):
            v = s.split("=")[1].strip().split()[0]
            res["elapsed"] = float(v)
    return res


def symmetry(e: np.ndarray, nx: int) -> dict:
    """Plane-0 energy symmetry measures, in the reference's loop order
    (reference/C/LULESH-OMP.cc:3203-3218)."""
    max_abs = 0.0
    total = 0.0
    max_rel = 0.0
    with np.errstate(invalid="ignore", divide="ignore"):
        for j in range(nx):
            for k in range(j + 1, nx):
                a = abs(e[j * nx + k] - e[k * nx + j])
                total += a
                if max_abs < a:
                    max_abs = a
                # 0/0 gives NaN, which never raises the maximum (as in C)
                r = a / e[k * nx + j]
                if max_rel < r:
                    max_rel = r
    return {"max_abs_diff": max_abs, "total_abs_diff": total,
            "max_rel_diff": max_rel}


def rand_sequence(seed: int, n: int) -> np.ndarray:
    out = np.empty(n, np.int32)
    lib().lo_rand_sequence(seed, n, iptr(out))
    return out
