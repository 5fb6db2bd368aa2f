"""Loader of the HIP step library (combustioncodegpu_amd/libcfx.so).

There is no fallback: if the library is missing or fails to load, every
entry point raises.  Build it with ``__graft_entry__.build()`` or
``python -m combustioncodegpu_amd.build``.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import abi
from .driver import StepContext

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcfx.so")
_lib = None

K_RHS, K_UPDATE, K_BC, K_PROPS, K_DT, K_OTHER = range(6)
KERNEL_CLASSES = ("rhs", "update", "ghost", "props", "dt", "other")


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
    lib = C.CDLL(LIB_PATH)
    abi.declare_api(lib, "cfx_")
    vp, dp = C.c_void_p, C.POINTER(C.c_double)
    lib.cfx_version.argtypes = []
    lib.cfx_version.restype = C.c_int
    lib.cfx_check.argtypes = [vp, C.POINTER(C.c_longlong)]
    lib.cfx_check.restype = C.c_int
    lib.cfx_sync.argtypes = [vp]
    lib.cfx_sync.restype = C.c_int
    lib.cfx_time_steps.argtypes = [vp, C.c_int, dp]
    lib.cfx_time_steps.restype = C.c_int
    lib.cfx_kernel_times.argtypes = [vp, C.c_int, dp, C.POINTER(C.c_longlong)]
    lib.cfx_kernel_times.restype = C.c_int
    lib.cfx_set_profiling.argtypes = [vp, C.c_int]
    lib.cfx_set_profiling.restype = C.c_int
    lib.cfx_rhs_variants.argtypes = [vp]
    lib.cfx_rhs_variants.restype = C.c_int
    lib.cfx_set_rhs_variant.argtypes = [vp, C.c_int]
    lib.cfx_set_rhs_variant.restype = C.c_int
    ip = C.POINTER(C.c_int)
    lib.cfx_set_rhs_tiles.argtypes = [vp, ip, ip]
    lib.cfx_set_rhs_tiles.restype = C.c_int
    lib.cfx_get_rhs_tiles.argtypes = [vp, ip, ip]
    lib.cfx_get_rhs_tiles.restype = C.c_int
    lib.cfx_autotune.argtypes = [vp, C.c_int, dp]
    lib.cfx_autotune.restype = C.c_int
    lib.cfx_set_execution.argtypes = [vp, C.c_int, C.c_int]
    lib.cfx_set_execution.restype = C.c_int
    lib.cfx_schedule_info.argtypes = [vp, ip, ip]
    lib.cfx_schedule_info.restype = C.c_int
    fp = C.POINTER(C.c_float)
    lib.cfx_mlp_default_config.argtypes = [C.POINTER(abi.CfxMlpConfig)]
    lib.cfx_mlp_default_config.restype = None
    lib.cfx_mlp_train.argtypes = [vp, C.POINTER(abi.CfxMlpConfig), C.POINTER(abi.CfxMlpReport)]
    lib.cfx_mlp_train.restype = C.c_int
    lib.cfx_mlp_n_params.argtypes = [vp]
    lib.cfx_mlp_n_params.restype = C.c_int
    lib.cfx_mlp_set_params.argtypes = [vp, fp, fp]
    lib.cfx_mlp_set_params.restype = C.c_int
    lib.cfx_mlp_get_params.argtypes = [vp, fp, fp]
    lib.cfx_mlp_get_params.restype = C.c_int
    lib.cfx_mlp_set_check.argtypes = [vp, C.c_double, C.c_int]
    lib.cfx_mlp_set_check.restype = C.c_int
    lib.cfx_mlp_gradient.argtypes = [vp, fp, fp, C.c_int, fp, dp]
    lib.cfx_mlp_gradient.restype = C.c_int
    lib.cfx_mlp_predict.argtypes = [vp, fp, C.c_int, fp]
    lib.cfx_mlp_predict.restype = C.c_int
    lib.cfx_mlp_predict_noalloc.argtypes = [vp, fp, C.c_int, fp]
    lib.cfx_mlp_predict_noalloc.restype = C.c_int
    lib.cfx_set_transport.argtypes = [vp, C.c_int]
    lib.cfx_set_transport.restype = C.c_int
    lib.cfx_mlp_stats.argtypes = [vp, C.POINTER(C.c_longlong)]
    lib.cfx_mlp_stats.restype = C.c_int
    lib.cfx_ghost_grid.argtypes = [C.POINTER(abi.CfxBlock), C.c_int, C.POINTER(dp)]
    lib.cfx_ghost_grid.restype = C.c_int
    lib.cfx_inflow_state.argtypes = [C.POINTER(abi.CfxParams), C.POINTER(abi.CfxSpecies), C.c_int, dp, dp]
    lib.cfx_inflow_state.restype = C.c_int
    lib.cfx_parse_numbers.argtypes = [C.c_char_p, C.c_longlong, dp, C.c_longlong, C.POINTER(C.c_longlong)]
    lib.cfx_parse_numbers.restype = C.c_int
    _lib = lib
    return lib


def ghost_grid(case):
    """Ghost-padded vertex coordinates of every block, (3, pz, py, px) each,
    from the library's host-side mesh setup (no device is used)."""
    lib = load()
    m = case.marshal()
    out = [np.empty((3,) + b.padded_shape, dtype=np.float64) for b in case.blocks]
    ptrs = (C.POINTER(C.c_double) * len(out))(*[abi.dptr(a) for a in out])
    rc = lib.cfx_ghost_grid(m.blocks, m.n_block, ptrs)
    if rc:
        raise RuntimeError(f"cfx_ghost_grid failed ({rc})")
    return out


def inflow_state(case, index):
    """(rho, u, v, w, p, T) and mass fractions of inflow record `index` as
    the library derives them (host only)."""
    lib = load()
    m = case.marshal()
    bv = np.empty(6, dtype=np.float64)
    yk = np.zeros(max(1, case.n_spec), dtype=np.float64)
    rc = lib.cfx_inflow_state(C.byref(m.params), C.byref(m.species), index, abi.dptr(bv), abi.dptr(yk))
    if rc:
        raise RuntimeError(f"cfx_inflow_state failed ({rc})")
    return bv, yk[:case.n_spec]


class Solver(StepContext):
    """One context of the HIP step library for a case."""

    def __init__(self, case):
        super().__init__(load(), "cfx_", case)

    def check(self):
        n = C.c_longlong(0)
        rc = self.lib.cfx_check(self.h, C.byref(n))
        return rc, n.value

    def sync(self):
        self._call("sync")

    def time_steps(self, n):
        ms = C.c_double(0.0)
        self._call("time_steps", int(n), C.byref(ms))
        return ms.value

    def set_profiling(self, on=True):
        self._call("set_profiling", 1 if on else 0)

    def rhs_variants(self):
        return self.lib.cfx_rhs_variants(self.h)

    def set_rhs_variant(self, v):
        self._call("set_rhs_variant", int(v))

    def set_rhs_tiles(self, variant, persist):
        v = (C.c_int * 3)(*variant)
        p = (C.c_int * 3)(*persist)
        self._call("set_rhs_tiles", v, p)

    def rhs_tiles(self):
        v = (C.c_int * 3)()
        p = (C.c_int * 3)()
        self._call("get_rhs_tiles", v, p)
        return list(v), list(p)

    def autotune(self, reps=3):
        """Pick the fastest RHS tile shape and persistence per direction on
        this context's blocks; returns (variants, persistence, ms per direction)."""
        ms = (C.c_double * 3)()
        self._call("autotune", int(reps), ms)
        v, p = self.rhs_tiles()
        return v, p, list(ms)

    def set_execution(self, batched=True, graph=False):
        self._call("set_execution", 1 if batched else 0, 1 if graph else 0)

    def schedule_info(self):
        n_ops = C.c_int(0)
        n_lev = C.c_int(0)
        self._call("schedule_info", C.byref(n_ops), C.byref(n_lev))
        return n_ops.value, n_lev.value

    # ------------------------------------------------ transport surrogate
    @staticmethod
    def mlp_default_config(**overrides):
        cfg = abi.CfxMlpConfig()
        load().cfx_mlp_default_config(C.byref(cfg))
        for k, v in overrides.items():
            setattr(cfg, k, v)
        return cfg

    def mlp_train(self, **overrides):
        """Train the surrogate on the GPU; returns the report as a dict."""
        cfg = self.mlp_default_config(**overrides)
        rep = abi.CfxMlpReport()
        self._call("mlp_train", C.byref(cfg), C.byref(rep))
        out = {k: getattr(rep, k) for k, _ in abi.CfxMlpReport._fields_}
        out["val_max_rel"] = list(rep.val_max_rel)
        out["val_p99_rel"] = list(rep.val_p99_rel)
        return out

    def mlp_n_params(self):
        return self.lib.cfx_mlp_n_params(self.h)

    def mlp_set_params(self, params, norm):
        p = np.ascontiguousarray(params, dtype=np.float32)
        n = np.ascontiguousarray(norm, dtype=np.float32)
        self._call("mlp_set_params", p.ctypes.data_as(C.POINTER(C.c_float)), n.ctypes.data_as(C.POINTER(C.c_float)))

    def mlp_get_params(self):
        p = np.zeros(self.mlp_n_params(), dtype=np.float32)
        n = np.zeros(64, dtype=np.float32)
        self._call("mlp_get_params", p.ctypes.data_as(C.POINTER(C.c_float)), n.ctypes.data_as(C.POINTER(C.c_float)))
        return p, n

    def mlp_set_check(self, tol, every):
        self._call("mlp_set_check", float(tol), int(every))

    def mlp_gradient(self, x, t):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        g = np.zeros(self.mlp_n_params(), dtype=np.float32)
        loss = C.c_double(0.0)
        f = C.POINTER(C.c_float)
        self._call("mlp_gradient", x.ctypes.data_as(f), t.ctypes.data_as(f), int(x.shape[0]), g.ctypes.data_as(f),
                   C.byref(loss))
        return g, loss.value

    def mlp_predict(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.zeros_like(x)
        f = C.POINTER(C.c_float)
        self._call("mlp_predict", x.ctypes.data_as(f), int(x.shape[0]), y.ctypes.data_as(f))
        return y

    def mlp_predict_noalloc(self, x, y=None):
        """cfx_mlp_predict_noalloc: y (n x 16 float32, written in place) from x."""
        x = np.ascontiguousarray(x, dtype=np.float32)
        if y is None:
            y = np.full_like(x, np.nan)
        f = C.POINTER(C.c_float)
        self._call("mlp_predict_noalloc", x.ctypes.data_as(f), int(x.shape[0]), y.ctypes.data_as(f))
        return y

    def set_transport(self, transport):
        self._call("set_transport", int(transport))

    def mlp_stats(self):
        c = (C.c_longlong * 3)()
        self._call("mlp_stats", c)
        return {"outside": c[0], "checked": c[1], "check_failed": c[2]}

    def kernel_times(self):
        out = {}
        for c, name in enumerate(KERNEL_CLASSES):
            ms = C.c_double(0.0)
            n = C.c_longlong(0)
            self.lib.cfx_kernel_times(self.h, c, C.byref(ms), C.byref(n))
            out[name] = (ms.value, n.value)
        return out


def parse_numbers(text: bytes, n_max: int):
    """Up to n_max whitespace-separated numbers of text (host only); raises
    ValueError naming the first token that is not a number."""
    lib = load()
    out = np.empty(max(0, n_max), dtype=np.float64)
    n = C.c_longlong(0)
    rc = lib.cfx_parse_numbers(text, len(text), abi.dptr(out), n_max, C.byref(n))
    if rc:
        raise ValueError(f"token {n.value} is not a number")
    return out[:n.value]
