"""Python binding of the MI355X LULESH step (the C ABI of
include/lulesh_mi355x.h), through ctypes.

This is the host-side mirror of the reference driver's view of the step:
`Domain` is the state the Accelerate program threads through `simulate`
(src/Main.hs:90-109, src/Domain.hs:11-21), and `step`/`run` are the
reference's cycle loop (reference/C/lulesh-2.0.3/lulesh.cc:2771-2781).
The library must be built (lulesh-accelerate_amd/build.py); there is no
fallback: without the HIP library every call raises.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "liblulesh_mi355x.so")

OK, VOLUME_ERROR, QSTOP_ERROR, BAD_ARG, HIP_ERROR = 0, -1, -2, -3, -4

NODE_FIELDS = {"x": 0, "y": 1, "z": 2, "xd": 3, "yd": 4, "zd": 5,
               "nodalMass": 6}
ELEM_FIELDS = {"e": 16, "p": 17, "q": 18, "ql": 19, "qq": 20, "v": 21,
               "volo": 22, "delv": 23, "vdov": 24, "arealg": 25, "ss": 26,
               "elemMass": 27, "vnew": 28, "delv_xi": 29, "delv_eta": 30,
               "delv_zeta": 31, "delx_xi": 32, "delx_eta": 33,
               "delx_zeta": 34}


class LuleshError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"LULESH error {code}: {msg}")
        self.code = code


class Config(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int), ("ic_mode", ctypes.c_int),
                ("num_reg", ctypes.c_int), ("balance", ctypes.c_int),
                ("cost", ctypes.c_int), ("max_cycles", ctypes.c_int),
                ("stoptime", ctypes.c_double)]


class Ensemble(ctypes.Structure):
    _fields_ = [("num_members", ctypes.c_int), ("nx_min", ctypes.c_int),
                ("nx_max", ctypes.c_int), ("efactor_min", ctypes.c_double),
                ("efactor_max", ctypes.c_double), ("seed", ctypes.c_ulonglong)]


class Stats(ctypes.Structure):
    _fields_ = [("time", ctypes.c_double), ("deltatime", ctypes.c_double),
                ("dtcourant", ctypes.c_double), ("dthydro", ctypes.c_double),
                ("e_origin", ctypes.c_double), ("cycle", ctypes.c_int),
                ("done", ctypes.c_int), ("num_elem", ctypes.c_int),
                ("num_node", ctypes.c_int)]


_lib = None

# every symbol declared in include/lulesh_mi355x.h
EXPORTS = (
    "lulesh_config_default", "lulesh_create", "lulesh_destroy",
    "lulesh_create_mixed", "lulesh_create_scaled", "lulesh_is_mixed", "lulesh_step",
    "lulesh_run", "lulesh_set_fused_multiply_add", "lulesh_upload_state",
    "lulesh_download_state",
    "lulesh_get_field", "lulesh_set_field", "lulesh_get_stats",
    "lulesh_set_scalars", "lulesh_set_parameter", "lulesh_set_sweep_shape",
    "lulesh_get_sweep_shape", "lulesh_get_regions", "lulesh_symmetry",
    "lulesh_kernel_count", "lulesh_kernel_name", "lulesh_set_kernel_timing",
    "lulesh_set_cycle_graphs",
    "lulesh_kernel_times", "lulesh_last_step_ms", "lulesh_last_error",
    "lulesh_device_math",
    "lulesh_ensemble_default", "lulesh_ensemble_draw", "lulesh_batch_create",
    "lulesh_batch_destroy", "lulesh_batch_step", "lulesh_batch_run",
    "lulesh_batch_get_stats", "lulesh_batch_get_field",
    "lulesh_batch_set_shape", "lulesh_batch_get_shape",
    "lulesh_batch_set_fused_multiply_add",
    "lulesh_batch_autotune", "lulesh_batch_last_ms",
    "lulesh_batch_last_error",
    "lulesh_mlp_param_count", "lulesh_mlp_create", "lulesh_mlp_destroy",
    "lulesh_mlp_get_params", "lulesh_mlp_eval", "lulesh_mlp_gradient",
    "lulesh_mlp_gradient_noalloc",
    "lulesh_mlp_train", "lulesh_mlp_last_error", "lulesh_set_learned_q",
    "lulesh_q_features", "lulesh_set_learned_q_form", "lulesh_write_domain_arrays",
    "lulesh_write_domain",
    "lulesh_create_lmesh", "lulesh_create_mesh", "lulesh_origin_element",
    "lulesh_get_nodelist", "lulesh_set_eos_bands", "lulesh_force_blocks",
)


def lib() -> ctypes.CDLL:
    """Loads the HIP library; raises when it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with "
            "`python lulesh-accelerate_amd/build.py` (no CPU fallback exists)")
    L = ctypes.CDLL(LIB_PATH)
    P, I, D = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
    DP = ctypes.POINTER(ctypes.c_double)
    IP = ctypes.POINTER(ctypes.c_int)
    L.lulesh_config_default.argtypes = [ctypes.POINTER(Config)]
    L.lulesh_config_default.restype = None
    L.lulesh_create.argtypes = [ctypes.POINTER(Config), ctypes.POINTER(P)]
    L.lulesh_create.restype = I
    L.lulesh_destroy.argtypes = [P]
    L.lulesh_destroy.restype = None
    L.lulesh_create_mixed.argtypes = [ctypes.POINTER(Config), ctypes.POINTER(P)]
    L.lulesh_create_mixed.restype = I
    L.lulesh_create_scaled.argtypes = [ctypes.POINTER(Config), D, I, ctypes.POINTER(P)]
    L.lulesh_create_scaled.restype = I
    L.lulesh_create_lmesh.argtypes = [ctypes.c_char_p, ctypes.POINTER(Config),
                                      ctypes.POINTER(P)]
    L.lulesh_create_lmesh.restype = I
    L.lulesh_create_mesh.argtypes = ([I, I, DP, DP, DP, IP, IP] + [I, IP] * 4 +
                                     [ctypes.POINTER(Config), ctypes.POINTER(P)])
    L.lulesh_create_mesh.restype = I
    L.lulesh_origin_element.argtypes = [P]
    L.lulesh_origin_element.restype = I
    L.lulesh_force_blocks.argtypes = [P, ctypes.POINTER(ctypes.c_longlong)]
    L.lulesh_force_blocks.restype = I
    L.lulesh_get_nodelist.argtypes = [P, IP]
    L.lulesh_get_nodelist.restype = I
    L.lulesh_set_eos_bands.argtypes = [P, I]
    L.lulesh_set_eos_bands.restype = I
    L.lulesh_is_mixed.argtypes = [P]
    L.lulesh_is_mixed.restype = I
    L.lulesh_step.argtypes = [P, I, IP]
    L.lulesh_step.restype = I
    L.lulesh_run.argtypes = [P, IP]
    L.lulesh_run.restype = I
    L.lulesh_set_fused_multiply_add.argtypes = [P, I]
    L.lulesh_set_fused_multiply_add.restype = I
    L.lulesh_upload_state.argtypes = [P] + [DP] * 11 + [D, D]
    L.lulesh_upload_state.restype = I
    L.lulesh_download_state.argtypes = [P] + [DP] * 13
    L.lulesh_download_state.restype = I
    L.lulesh_get_field.argtypes = [P, I, DP]
    L.lulesh_get_field.restype = I
    L.lulesh_set_field.argtypes = [P, I, DP]
    L.lulesh_set_field.restype = I
    L.lulesh_get_stats.argtypes = [P, ctypes.POINTER(Stats)]
    L.lulesh_get_stats.restype = I
    L.lulesh_set_scalars.argtypes = [P, DP, DP, DP, DP, IP]
    L.lulesh_set_scalars.restype = I
    L.lulesh_set_parameter.argtypes = [P, ctypes.c_char_p, D]
    L.lulesh_set_parameter.restype = I
    L.lulesh_set_sweep_shape.argtypes = [P, I, I, I, I]
    L.lulesh_set_sweep_shape.restype = I
    L.lulesh_get_sweep_shape.argtypes = [P, IP, IP, IP, IP]
    L.lulesh_get_sweep_shape.restype = I
    L.lulesh_get_regions.argtypes = [P, IP]
    L.lulesh_get_regions.restype = I
    L.lulesh_symmetry.argtypes = [P, DP, DP, DP]
    L.lulesh_symmetry.restype = I
    L.lulesh_kernel_count.argtypes = []
    L.lulesh_kernel_count.restype = I
    L.lulesh_kernel_name.argtypes = [I]
    L.lulesh_kernel_name.restype = ctypes.c_char_p
    L.lulesh_set_kernel_timing.argtypes = [P, I]
    L.lulesh_set_kernel_timing.restype = I
    L.lulesh_set_cycle_graphs.argtypes = [P, I]
    L.lulesh_set_cycle_graphs.restype = I
    L.lulesh_kernel_times.argtypes = [P, DP, ctypes.POINTER(ctypes.c_longlong)]
    L.lulesh_kernel_times.restype = I
    L.lulesh_last_step_ms.argtypes = [P]
    L.lulesh_last_step_ms.restype = D
    L.lulesh_last_error.argtypes = [P]
    L.lulesh_last_error.restype = ctypes.c_char_p
    L.lulesh_device_math.argtypes = [I, I, DP, DP, DP]
    L.lulesh_device_math.restype = I
    L.lulesh_ensemble_default.argtypes = [ctypes.POINTER(Ensemble)]
    L.lulesh_ensemble_default.restype = None
    L.lulesh_ensemble_draw.argtypes = [ctypes.POINTER(Ensemble), IP, DP]
    L.lulesh_ensemble_draw.restype = I
    L.lulesh_batch_create.argtypes = [ctypes.POINTER(Config), DP, I,
                                      ctypes.POINTER(P)]
    L.lulesh_batch_create.restype = I
    L.lulesh_batch_destroy.argtypes = [P]
    L.lulesh_batch_destroy.restype = None
    L.lulesh_batch_step.argtypes = [P, I, IP]
    L.lulesh_batch_step.restype = I
    L.lulesh_batch_run.argtypes = [P]
    L.lulesh_batch_run.restype = I
    L.lulesh_batch_get_stats.argtypes = [P, I, ctypes.POINTER(Stats)]
    L.lulesh_batch_get_stats.restype = I
    L.lulesh_batch_get_field.argtypes = [P, I, I, DP]
    L.lulesh_batch_get_field.restype = I
    L.lulesh_batch_set_fused_multiply_add.argtypes = [P, I]
    L.lulesh_batch_set_fused_multiply_add.restype = I
    L.lulesh_batch_set_shape.argtypes = [P] + [I] * 7
    L.lulesh_batch_set_shape.restype = I
    L.lulesh_batch_get_shape.argtypes = [P] + [IP] * 7
    L.lulesh_batch_get_shape.restype = I
    L.lulesh_batch_autotune.argtypes = [P, I, DP]
    L.lulesh_batch_autotune.restype = I
    L.lulesh_batch_last_ms.argtypes = [P]
    L.lulesh_batch_last_ms.restype = D
    L.lulesh_batch_last_error.argtypes = [P]
    L.lulesh_batch_last_error.restype = ctypes.c_char_p
    FP = ctypes.POINTER(ctypes.c_float)
    L.lulesh_mlp_param_count.argtypes = []
    L.lulesh_mlp_param_count.restype = I
    L.lulesh_mlp_create.argtypes = [FP, FP, FP, ctypes.POINTER(P)]
    L.lulesh_mlp_create.restype = I
    L.lulesh_mlp_destroy.argtypes = [P]
    L.lulesh_mlp_destroy.restype = None
    L.lulesh_mlp_get_params.argtypes = [P, FP]
    L.lulesh_mlp_get_params.restype = I
    L.lulesh_mlp_eval.argtypes = [P, FP, I, FP]
    L.lulesh_mlp_eval.restype = I
    L.lulesh_mlp_gradient.argtypes = [P, FP, FP, I, FP, FP, DP]
    L.lulesh_mlp_gradient.restype = I
    L.lulesh_mlp_gradient_noalloc.argtypes = [P, P, P, I, FP, P, P, P]
    L.lulesh_mlp_gradient_noalloc.restype = I
    L.lulesh_mlp_train.argtypes = [P, FP, FP, I, FP, I, I, ctypes.c_float,
                                   ctypes.c_ulonglong, DP]
    L.lulesh_mlp_train.restype = I
    L.lulesh_mlp_last_error.argtypes = [P]
    L.lulesh_mlp_last_error.restype = ctypes.c_char_p
    L.lulesh_set_learned_q.argtypes = [P, P, D]
    L.lulesh_set_learned_q.restype = I
    L.lulesh_q_features.argtypes = [P, FP, DP]
    L.lulesh_q_features.restype = I
    L.lulesh_set_learned_q_form.argtypes = [P, I]
    L.lulesh_set_learned_q_form.restype = I
    L.lulesh_write_domain_arrays.argtypes = [ctypes.c_char_p, I, I, D] + [DP] * 10
    L.lulesh_write_domain_arrays.restype = I
    L.lulesh_write_domain.argtypes = [P, ctypes.c_char_p]
    L.lulesh_write_domain.restype = I
    _lib = L
    return L


def _dp(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def default_config() -> Config:
    c = Config()
    lib().lulesh_config_default(ctypes.byref(c))
    return c


def _apply_cycle_graphs_env(d) -> None:
    """LULESH_CYCLE_GRAPHS=0 in the environment turns the captured cycle
    graphs off for every domain this module creates (for tools that trace
    kernel launches one by one)."""
    if os.environ.get("LULESH_CYCLE_GRAPHS", "1") == "0":
        d.set_cycle_graphs(False)


class Domain:
    """A Sedov cube of nx^3 elements resident on the GPU."""

    def __init__(self, nx: int = 30, ic_mode: int = 2, num_reg: int = 11,
                 balance: int = 1, cost: int = 1, max_cycles: int = 9999999,
                 stoptime: float = 1.0e-2, mixed: bool = False,
                 efactor: float = 1.0, fused: bool = False):
        """mixed=True builds the mixed-precision form (lulesh_create_mixed);
        efactor multiplies the initial origin energy; fused=True runs the
        float64 step with fused multiply-adds
        (lulesh_set_fused_multiply_add)."""
        L = lib()
        cfg = default_config()
        cfg.nx, cfg.ic_mode, cfg.num_reg = nx, ic_mode, num_reg
        cfg.balance, cfg.cost = balance, cost
        cfg.max_cycles, cfg.stoptime = max_cycles, stoptime
        h = ctypes.c_void_p()
        if efactor != 1.0:
            rc = L.lulesh_create_scaled(ctypes.byref(cfg), float(efactor), int(mixed),
                                        ctypes.byref(h))
        elif mixed:
            rc = L.lulesh_create_mixed(ctypes.byref(cfg), ctypes.byref(h))
        else:
            rc = L.lulesh_create(ctypes.byref(cfg), ctypes.byref(h))
        if rc != OK:
            raise LuleshError(rc, "lulesh_create failed")
        self.h = h
        self.nx, self.ic_mode = nx, ic_mode
        self.mixed = bool(L.lulesh_is_mixed(h))
        self.fused = False
        if fused:
            self.set_fused_multiply_add(True)
        _apply_cycle_graphs_env(self)
        st = self.stats()
        self.num_elem, self.num_node = st.num_elem, st.num_node

    def set_fused_multiply_add(self, enable: bool) -> None:
        self._check(lib().lulesh_set_fused_multiply_add(self.h, int(bool(enable))))
        self.fused = bool(enable)

    @classmethod
    def from_lmesh(cls, path: str, stoptime: float = 1.0e-2,
                   max_cycles: int = 9999999, fused: bool = False) -> "Domain":
        """An unstructured mesh read from a .lmesh file (LULESH 1.0
        semantics); fused as for the cube."""
        L = lib()
        cfg = default_config()
        cfg.ic_mode, cfg.stoptime, cfg.max_cycles = 1, stoptime, max_cycles
        h = ctypes.c_void_p()
        rc = L.lulesh_create_lmesh(path.encode(), ctypes.byref(cfg), ctypes.byref(h))
        if rc != OK:
            raise LuleshError(rc, f"lulesh_create_lmesh failed for {path}")
        return cls._wrap(h, 1, fused)

    @classmethod
    def from_mesh(cls, mesh: dict, stoptime: float = 1.0e-2,
                  max_cycles: int = 9999999, fused: bool = False) -> "Domain":
        """An unstructured mesh from arrays: x, y, z, nodelist [ne, 8],
        neighbors [ne, 6], symm_x, symm_y, symm_z, free_surf; fused as for
        the cube."""
        L = lib()
        cfg = default_config()
        cfg.ic_mode, cfg.stoptime, cfg.max_cycles = 1, stoptime, max_cycles
        a = {k: np.ascontiguousarray(mesh[k], np.float64) for k in ("x", "y", "z")}
        ints = {k: np.ascontiguousarray(mesh[k], np.int32)
                for k in ("nodelist", "neighbors", "symm_x", "symm_y", "symm_z",
                          "free_surf")}
        ip = lambda v: v.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
        h = ctypes.c_void_p()
        rc = L.lulesh_create_mesh(
            len(ints["nodelist"].reshape(-1, 8)), len(a["x"]), _dp(a["x"]), _dp(a["y"]),
            _dp(a["z"]), ip(ints["nodelist"]), ip(ints["neighbors"]),
            len(ints["symm_x"]), ip(ints["symm_x"]), len(ints["symm_y"]), ip(ints["symm_y"]),
            len(ints["symm_z"]), ip(ints["symm_z"]), len(ints["free_surf"]),
            ip(ints["free_surf"]), ctypes.byref(cfg), ctypes.byref(h))
        if rc != OK:
            raise LuleshError(rc, "lulesh_create_mesh failed")
        return cls._wrap(h, 1, fused)

    @classmethod
    def _wrap(cls, h, ic_mode: int, fused: bool = False) -> "Domain":
        d = cls.__new__(cls)
        d.h = h
        d.nx, d.ic_mode, d.mixed, d.fused = 0, ic_mode, False, False
        if fused:
            d.set_fused_multiply_add(True)
        _apply_cycle_graphs_env(d)
        st = d.stats()
        d.num_elem, d.num_node = st.num_elem, st.num_node
        d.origin = lib().lulesh_origin_element(h)
        return d

    def close(self) -> None:
        if getattr(self, "h", None):
            lib().lulesh_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc: int) -> int:
        if rc != OK:
            msg = lib().lulesh_last_error(self.h).decode()
            raise LuleshError(rc, msg)
        return rc

    def step(self, ncycles: int = 1, raise_on_error: bool = True) -> tuple[int, int]:
        """Advances up to ncycles cycles; returns (rc, cycles done)."""
        n = ctypes.c_int(0)
        rc = lib().lulesh_step(self.h, int(ncycles), ctypes.byref(n))
        if raise_on_error:
            self._check(rc)
        return rc, n.value

    def run(self, raise_on_error: bool = True) -> tuple[int, int]:
        n = ctypes.c_int(0)
        rc = lib().lulesh_run(self.h, ctypes.byref(n))
        if raise_on_error:
            self._check(rc)
        return rc, n.value

    def get(self, name: str) -> np.ndarray:
        if name in NODE_FIELDS:
            fid, n = NODE_FIELDS[name], self.num_node
        else:
            fid, n = ELEM_FIELDS[name], self.num_elem
        out = np.empty(n, np.float64)
        self._check(lib().lulesh_get_field(self.h, fid, _dp(out)))
        return out

    def set(self, name: str, values) -> None:
        if name in NODE_FIELDS:
            fid, n = NODE_FIELDS[name], self.num_node
        else:
            fid, n = ELEM_FIELDS[name], self.num_elem
        v = np.ascontiguousarray(values, np.float64)
        if v.size != n:
            raise ValueError(f"{name}: expected {n} values, got {v.size}")
        self._check(lib().lulesh_set_field(self.h, fid, _dp(v)))

    def stats(self) -> Stats:
        s = Stats()
        self._check(lib().lulesh_get_stats(self.h, ctypes.byref(s)))
        return s

    def set_scalars(self, time=None, deltatime=None, dtcourant=None,
                    dthydro=None, cycle=None) -> None:
        def dref(v):
            return ctypes.byref(ctypes.c_double(v)) if v is not None else None
        c = ctypes.byref(ctypes.c_int(int(cycle))) if cycle is not None else None
        self._check(lib().lulesh_set_scalars(self.h, dref(time), dref(deltatime),
                                             dref(dtcourant), dref(dthydro), c))

    def set_parameter(self, name: str, value: float) -> None:
        """Sets a constant of the method (hgcoef, qstop, ..., stoptime)."""
        self._check(lib().lulesh_set_parameter(self.h, name.encode(),
                                               float(value)))

    def set_sweep_shape(self, tx: int, ty: int, nz: int, waves: int) -> None:
        self._check(lib().lulesh_set_sweep_shape(self.h, tx, ty, nz, waves))

    def set_eos_bands(self, bands: int) -> None:
        """Interleave of the Q/EOS schedule (lulesh_set_eos_bands)."""
        self._check(lib().lulesh_set_eos_bands(self.h, int(bands)))

    def sweep_shape(self) -> tuple[int, int, int, int]:
        v = [ctypes.c_int() for _ in range(4)]
        self._check(lib().lulesh_get_sweep_shape(self.h, *[ctypes.byref(a) for a in v]))
        return tuple(a.value for a in v)

    def regions(self) -> np.ndarray:
        out = np.empty(self.num_elem, np.int32)
        self._check(lib().lulesh_get_regions(
            self.h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int))))
        return out

    def symmetry(self) -> dict:
        a, t, r = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        self._check(lib().lulesh_symmetry(self.h, ctypes.byref(a),
                                          ctypes.byref(t), ctypes.byref(r)))
        return {"max_abs_diff": a.value, "total_abs_diff": t.value,
                "max_rel_diff": r.value}

    def download_state(self) -> dict:
        nn, ne = self.num_node, self.num_elem
        arr = {k: np.empty(nn) for k in ("x", "y", "z", "xd", "yd", "zd")}
        arr.update({k: np.empty(ne) for k in ("e", "p", "q", "v", "ss")})
        t, dt = ctypes.c_double(), ctypes.c_double()
        order = ("x", "y", "z", "xd", "yd", "zd", "e", "p", "q", "v", "ss")
        self._check(lib().lulesh_download_state(
            self.h, *[_dp(arr[k]) for k in order], ctypes.byref(t),
            ctypes.byref(dt)))
        arr["time"], arr["deltatime"] = t.value, dt.value
        return arr

    def upload_state(self, st: dict) -> None:
        order = ("x", "y", "z", "xd", "yd", "zd", "e", "p", "q", "v", "ss")
        bufs = [np.ascontiguousarray(st[k], np.float64) for k in order]
        self._check(lib().lulesh_upload_state(
            self.h, *[_dp(b) for b in bufs], float(st["time"]),
            float(st["deltatime"])))

    # -- learned closure ------------------------------------------------
    LEARNED_FACTORS, LEARNED_STANDIN = 0, 1

    def set_learned_q(self, net, z_max: float = 4.0, form: int | None = None) -> None:
        """Uses the network `net` (an Mlp, kept alive by this domain) in the
        Q phase, in the given form (LEARNED_FACTORS: factors on the
        limiter's Q; LEARNED_STANDIN: in the limiter's place; None: keep
        the form set); net None restores the limiter."""
        if form is not None:
            self.set_learned_q_form(form)
        self._net = net
        self._check(lib().lulesh_set_learned_q(self.h, net.h if net else None,
                                               float(z_max)))

    def set_learned_q_form(self, form: int) -> None:
        self._check(lib().lulesh_set_learned_q_form(self.h, int(form)))

    def q_features(self) -> tuple[np.ndarray, np.ndarray]:
        """(features[num_elem, 12], q[num_elem, 4]) of the current state."""
        f = np.empty((self.num_elem, 12), np.float32)
        q = np.empty((self.num_elem, 4), np.float64)
        rc = lib().lulesh_q_features(self.h, _fp(f), _dp(q))
        if rc not in (OK, VOLUME_ERROR):
            self._check(rc)
        return f, q

    def nodelist(self) -> np.ndarray:
        """The element corners, [num_elem, 8] node indices."""
        out = np.empty((self.stats().num_elem, 8), np.int32)
        self._check(lib().lulesh_get_nodelist(
            self.h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int))))
        return out

    def write(self, path: str) -> None:
        """The state as a VTK file (lulesh_write_domain)."""
        self._check(lib().lulesh_write_domain(self.h, path.encode()))

    # -- integral diagnostics -------------------------------------------
    def total_energy(self) -> float:
        """Internal plus kinetic energy: sum(e * elemMass) +
        1/2 sum(nodalMass * |xd|^2)."""
        ie = float(np.dot(self.get("e"), self.get("elemMass")))
        v2 = self.get("xd") ** 2 + self.get("yd") ** 2 + self.get("zd") ** 2
        return ie + 0.5 * float(np.dot(self.get("nodalMass"), v2))

    def force_blocks(self) -> dict:
        """How the unstructured force kernel cuts the mesh
        (lulesh_force_blocks): element blocks, interior nodes (summed on
        chip), boundary nodes and their element corners (through device
        memory)."""
        out = (ctypes.c_longlong * 4)()
        self._check(lib().lulesh_force_blocks(self.h, out))
        return {"blocks": out[0], "interior_nodes": out[1], "boundary_nodes": out[2],
                "boundary_corners": out[3]}

    def shock_radius(self) -> dict:
        """Distance from the origin of the centroid of the element with the
        largest pressure, over the whole mesh and along the x axis."""
        n, en = self.nx, self.nx + 1
        p = self.get("p").reshape(n, n, n)          # [k, j, i]
        x = self.get("x").reshape(en, en, en)
        y = self.get("y").reshape(en, en, en)
        z = self.get("z").reshape(en, en, en)

        def centroid(a):
            return 0.125 * (a[:-1, :-1, :-1] + a[:-1, :-1, 1:] + a[:-1, 1:, :-1] +
                            a[:-1, 1:, 1:] + a[1:, :-1, :-1] + a[1:, :-1, 1:] +
                            a[1:, 1:, :-1] + a[1:, 1:, 1:])
        r = np.sqrt(centroid(x) ** 2 + centroid(y) ** 2 + centroid(z) ** 2)
        k = int(np.argmax(p))
        ix = int(np.argmax(p[0, 0, :]))
        return {"radius": float(r.reshape(-1)[k]),
                "radius_x_axis": float(r[0, 0, ix])}

    # -- timing ---------------------------------------------------------
    def set_cycle_graphs(self, enable: bool) -> None:
        """Cycles replayed from captured graphs (default) or queued one
        launch at a time (lulesh_set_cycle_graphs); the same bits."""
        self._check(lib().lulesh_set_cycle_graphs(self.h, int(bool(enable))))

    def set_kernel_timing(self, enable: bool) -> None:
        self._check(lib().lulesh_set_kernel_timing(self.h, int(bool(enable))))

    def kernel_times(self) -> dict:
        k = lib().lulesh_kernel_count()
        ms = (ctypes.c_double * k)()
        n = (ctypes.c_longlong * k)()
        self._check(lib().lulesh_kernel_times(self.h, ms, n))
        return {lib().lulesh_kernel_name(i).decode(): (ms[i], n[i])
                for i in range(k)}

    def last_step_ms(self) -> float:
        return lib().lulesh_last_step_ms(self.h)


def device_math(op: str, a, b=None) -> np.ndarray:
    """Evaluates the library's cube root, sqrt, division, the EOS's square
    root (inputs >= 2^-767), or the fused build's a / b or a / cbrt(b)
    (b > 0 normal) on the GPU."""
    code = {"cbrt": 0, "sqrt": 1, "div": 2, "sqrt_eos": 3, "div_fused": 4,
            "over_cbrt_fused": 5}[op]
    a = np.ascontiguousarray(a, np.float64)
    bb = np.ascontiguousarray(b, np.float64) if b is not None else None
    out = np.empty_like(a)
    rc = lib().lulesh_device_math(code, a.size, _dp(a),
                                  _dp(bb) if bb is not None else None, _dp(out))
    if rc != OK:
        raise LuleshError(rc, "lulesh_device_math failed")
    return out


def draw_ensemble(num_members: int = 64, nx_min: int = 16, nx_max: int = 64,
                  efactor_min: float = 0.8, efactor_max: float = 1.2,
                  seed: int = 0) -> tuple[np.ndarray, np.ndarray]:
    """Edge sizes and origin-energy factors of a seeded ensemble (the
    library's splitmix64 stream)."""
    e = Ensemble()
    lib().lulesh_ensemble_default(ctypes.byref(e))
    e.num_members, e.nx_min, e.nx_max = num_members, nx_min, nx_max
    e.efactor_min, e.efactor_max, e.seed = efactor_min, efactor_max, seed
    nx = np.empty(num_members, np.int32)
    ef = np.empty(num_members, np.float64)
    rc = lib().lulesh_ensemble_draw(
        ctypes.byref(e), nx.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), _dp(ef))
    if rc != OK:
        raise LuleshError(rc, "lulesh_ensemble_draw failed")
    return nx, ef


class Batch:
    """An ensemble of independent Sedov domains advanced together.

    Member i is the domain Domain(nx[i], ...) would build, with its initial
    origin energy multiplied by efactor[i]; every member keeps its own time,
    dt and stop test.
    """

    def __init__(self, nx, efactor=None, ic_mode: int = 2, num_reg: int = 11,
                 balance: int = 1, cost: int = 1, max_cycles: int = 9999999,
                 stoptime: float = 1.0e-2, fused: bool = False):
        """fused=True runs the phases with fused multiply-adds
        (lulesh_batch_set_fused_multiply_add)."""
        L = lib()
        nx = [int(v) for v in nx]
        n = len(nx)
        cfgs = (Config * n)()
        for i, v in enumerate(nx):
            L.lulesh_config_default(ctypes.byref(cfgs[i]))
            c = cfgs[i]
            c.nx, c.ic_mode, c.num_reg, c.balance, c.cost = v, ic_mode, num_reg, balance, cost
            c.max_cycles, c.stoptime = max_cycles, stoptime
        ef = None
        if efactor is not None:
            self._ef = np.ascontiguousarray(efactor, np.float64)
            if self._ef.size != n:
                raise ValueError("efactor needs one value per member")
            ef = _dp(self._ef)
        h = ctypes.c_void_p()
        rc = L.lulesh_batch_create(cfgs, ef, n, ctypes.byref(h))
        if rc != OK:
            raise LuleshError(rc, "lulesh_batch_create failed")
        self.h = h
        self.nx = np.array(nx)
        self.n = n
        if fused:
            self._check(L.lulesh_batch_set_fused_multiply_add(h, 1))

    def close(self) -> None:
        if getattr(self, "h", None):
            lib().lulesh_batch_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc: int) -> int:
        if rc != OK:
            raise LuleshError(rc, lib().lulesh_batch_last_error(self.h).decode())
        return rc

    def step(self, ncycles: int) -> int:
        """Advances every unfinished member by up to ncycles cycles; returns
        the number of members with cycles left."""
        p = ctypes.c_int(0)
        self._check(lib().lulesh_batch_step(self.h, int(ncycles), ctypes.byref(p)))
        return p.value

    def run(self, raise_on_error: bool = True) -> int:
        rc = lib().lulesh_batch_run(self.h)
        if raise_on_error:
            self._check(rc)
        return rc

    def stats(self, member: int) -> tuple[int, Stats]:
        """(error code of the member, its scalars)."""
        s = Stats()
        rc = lib().lulesh_batch_get_stats(self.h, int(member), ctypes.byref(s))
        if rc == BAD_ARG:
            self._check(rc)
        return rc, s

    def get(self, member: int, name: str) -> np.ndarray:
        _, st = self.stats(member)
        if name in NODE_FIELDS:
            fid, n = NODE_FIELDS[name], st.num_node
        else:
            fid, n = ELEM_FIELDS[name], st.num_elem
        out = np.empty(n, np.float64)
        self._check(lib().lulesh_batch_get_field(self.h, int(member), fid, _dp(out)))
        return out

    def set_shape(self, tx: int, ty: int, nz: int, waves: int, sweep_blocks: int = 0,
                  elem_blocks: int = 0, cycles_per_graph: int = 16) -> None:
        self._check(lib().lulesh_batch_set_shape(self.h, tx, ty, nz, waves,
                                                 sweep_blocks, elem_blocks,
                                                 cycles_per_graph))

    def shape(self) -> dict:
        v = [ctypes.c_int() for _ in range(7)]
        self._check(lib().lulesh_batch_get_shape(self.h, *[ctypes.byref(a) for a in v]))
        keys = ("tx", "ty", "nz", "waves", "sweep_blocks", "elem_blocks",
                "cycles_per_graph")
        return dict(zip(keys, (a.value for a in v)))

    def autotune(self, trial_cycles: int = 16) -> float:
        """Picks the fastest launch shape; returns its ms per cycle."""
        ms = ctypes.c_double(0.0)
        self._check(lib().lulesh_batch_autotune(self.h, int(trial_cycles),
                                                ctypes.byref(ms)))
        return ms.value

    def last_ms(self) -> float:
        return lib().lulesh_batch_last_ms(self.h)

    def zone_cycles(self) -> int:
        """Sum over members of elements x cycles done."""
        tot = 0
        for i in range(self.n):
            _, st = self.stats(i)
            tot += int(st.num_elem) * int(st.cycle)
        return tot


def _fp(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


class Mlp:
    """The learned-Q network on the GPU (12-64-64-2, bf16 MFMA, fp32
    master weights)."""

    N_IN, N_OUT = 12, 2

    def __init__(self, params, in_mean, in_std):
        L = lib()
        self.n_params = L.lulesh_mlp_param_count()
        p = np.ascontiguousarray(params, np.float32)
        m = np.ascontiguousarray(in_mean, np.float32)
        sd = np.ascontiguousarray(in_std, np.float32)
        if p.size != self.n_params or m.size != 12 or sd.size != 12:
            raise ValueError("parameter or normalisation size mismatch")
        h = ctypes.c_void_p()
        rc = L.lulesh_mlp_create(_fp(p), _fp(m), _fp(sd), ctypes.byref(h))
        if rc != OK:
            raise LuleshError(rc, "lulesh_mlp_create failed")
        self.h = h
        self.in_mean, self.in_std = m.copy(), sd.copy()

    def close(self) -> None:
        if getattr(self, "h", None):
            lib().lulesh_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc: int) -> None:
        if rc != OK:
            raise LuleshError(rc, lib().lulesh_mlp_last_error(self.h).decode())

    def params(self) -> np.ndarray:
        out = np.empty(self.n_params, np.float32)
        self._check(lib().lulesh_mlp_get_params(self.h, _fp(out)))
        return out

    def eval(self, x) -> np.ndarray:
        x = np.ascontiguousarray(x, np.float32).reshape(-1, 12)
        y = np.empty((len(x), 2), np.float32)
        self._check(lib().lulesh_mlp_eval(self.h, _fp(x), len(x), _fp(y)))
        return y

    def gradient(self, x, t, out_weight=(1.0, 1.0)) -> tuple[np.ndarray, float]:
        x = np.ascontiguousarray(x, np.float32).reshape(-1, 12)
        t = np.ascontiguousarray(t, np.float32).reshape(-1, 2)
        w = np.ascontiguousarray(out_weight, np.float32)
        g = np.empty(self.n_params, np.float32)
        loss = ctypes.c_double()
        self._check(lib().lulesh_mlp_gradient(self.h, _fp(x), _fp(t), len(x), _fp(w),
                                              _fp(g), ctypes.byref(loss)))
        return g, loss.value

    def gradient_noalloc(self, x_dev: int, t_dev: int, count: int, grad_dev: int,
                         loss_dev: int, stream: int = 0,
                         out_weight=(1.0, 1.0)) -> None:
        """lulesh_mlp_gradient_noalloc on device buffers given by address
        (x[count][12], t[count][2], grad[params], loss[1], float32), queued
        on `stream` (a hipStream_t handle, 0 = the null stream)."""
        w = np.ascontiguousarray(out_weight, np.float32)
        self._check(lib().lulesh_mlp_gradient_noalloc(
            self.h, x_dev, t_dev, int(count), _fp(w), grad_dev, loss_dev,
            stream or None))

    def train(self, x, t, epochs: int, batch: int = 1024, lr: float = 1e-3,
              seed: int = 0, out_weight=(1.0, 1.0)) -> float:
        x = np.ascontiguousarray(x, np.float32).reshape(-1, 12)
        t = np.ascontiguousarray(t, np.float32).reshape(-1, 2)
        w = np.ascontiguousarray(out_weight, np.float32)
        loss = ctypes.c_double()
        self._check(lib().lulesh_mlp_train(self.h, _fp(x), _fp(t), len(x), _fp(w),
                                           int(epochs), int(batch), float(lr), int(seed),
                                           ctypes.byref(loss)))
        return loss.value


def write_domain_arrays(path: str, step: int, time: float, st: dict) -> None:
    """writeDomain on host arrays (x, y, z, xd, yd, zd, e, p, v, q)."""
    arrs = [np.ascontiguousarray(st[k], np.float64)
            for k in ("x", "y", "z", "xd", "yd", "zd", "e", "p", "v", "q")]
    rc = lib().lulesh_write_domain_arrays(path.encode(), len(arrs[6]), int(step),
                                          float(time), *[_dp(a) for a in arrs])
    if rc != OK:
        raise LuleshError(rc, "lulesh_write_domain_arrays failed")
