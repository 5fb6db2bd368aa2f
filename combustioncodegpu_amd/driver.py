"""Host-side mirror of the reference's Driver around a step library.

`StepContext` wraps one context of a library that exports the step API of
include/cfx.h under a symbol prefix (``cfx_`` for the HIP library).  Its
methods are the pieces of Driver::steady_simulation (Driver.cu:113-192):
store_last_step, rhs, time_step, update, apply_bc, exchange,
eliminate_k_gradient, update_properties and residual, plus `step` for the
whole loop body, so callers read like the reference's own step loop.

`Driver` adds the reference's residual bookkeeping (Driver.cu:219-277): the
four RMS residuals are normalised by their value at the first output step
and the maximum is reported.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import abi


class CfxError(RuntimeError):
    def __init__(self, rc, msg):
        super().__init__(f"{msg} (rc={rc})")
        self.rc = rc


class StepContext:
    def __init__(self, lib, prefix: str, case):
        self.lib = lib
        self.prefix = prefix
        self.case = case
        self._m = case.marshal()
        h = C.c_void_p()
        rc = getattr(lib, prefix + "create")(C.byref(self._m.params), C.byref(self._m.species), self._m.blocks,
                                            self._m.n_block, C.byref(h))
        if rc != 0:
            raise CfxError(rc, f"{prefix}create failed")
        self.h = h

    # -- plumbing --------------------------------------------------------
    def _call(self, name, *args):
        rc = getattr(self.lib, self.prefix + name)(self.h, *args)
        if rc != 0:
            err = getattr(self.lib, self.prefix + "last_error")(self.h)
            raise CfxError(rc, f"{self.prefix}{name}: {err.decode() if err else ''}")
        return rc

    def close(self):
        if getattr(self, "h", None):
            getattr(self.lib, self.prefix + "destroy")(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- state transfer --------------------------------------------------
    def upload_primitive(self, blk, bv, yk=None):
        bv = np.ascontiguousarray(bv, dtype=np.float64)
        ykp = None
        if yk is not None:
            yk = np.ascontiguousarray(yk, dtype=np.float64)
            ykp = abi.dptr(yk)
        self._call("upload_primitive", blk, abi.dptr(bv), ykp)

    def upload_state(self, blk, cv, bv, yk=None):
        cv = np.ascontiguousarray(cv, dtype=np.float64)
        bv = np.ascontiguousarray(bv, dtype=np.float64)
        ykp = None
        if yk is not None:
            yk = np.ascontiguousarray(yk, dtype=np.float64)
            ykp = abi.dptr(yk)
        self._call("upload_state", blk, abi.dptr(cv), abi.dptr(bv), ykp)

    def upload_initial(self):
        for b in range(len(self.case.blocks)):
            self.upload_primitive(b, self.case.bv0[b], self.case.yk0[b])

    def shape_of(self, blk, what):
        B = self.case.blocks[blk]
        pad = B.padded_shape
        inner = (B.mz, B.my, B.mx)
        ns, nv = self.case.n_spec, self.case.n_var
        return {
            abi.CV: (nv,) + pad, abi.BV: (6,) + pad, abi.YK: (ns,) + pad, abi.PROPS: (4 + ns,) + pad,
            abi.DQ: (nv,) + inner, abi.METRIC: (10,) + pad, abi.DTLOCAL: inner, abi.COORD: (3,) + pad,
        }[what]

    def download(self, blk, what):
        out = np.empty(self.shape_of(blk, what), dtype=np.float64)
        self._call("download", blk, what, abi.dptr(out))
        return out

    # -- the step pieces -------------------------------------------------
    def init_conservative(self):
        self._call("init_conservative")

    def store_last_step(self):
        self._call("store_last_step")

    def rhs(self, blk=-1):
        self._call("rhs", blk)

    def time_step(self, global_=True):
        dt = C.c_double(0.0)
        self._call("time_step", 1 if global_ else 0, C.byref(dt))
        return dt.value

    def update(self, stage=0):
        self._call("update", stage)

    def lusgs(self):
        self._call("lusgs")

    def apply_bc(self):
        self._call("apply_bc")

    def exchange(self):
        self._call("exchange")

    def eliminate_k_gradient(self):
        self._call("eliminate_k_gradient")

    def update_properties(self):
        self._call("update_properties")

    def residual(self):
        r = (C.c_double * 4)()
        self._call("residual", r)
        return np.array(r[:])

    def step(self, n=1):
        self._call("step", n)

    def get_dt(self):
        dt = C.c_double(0.0)
        self._call("get_dt", C.byref(dt))
        return dt.value

    def set_dt(self, dt):
        self._call("set_dt", float(dt))


class Driver:
    """Driver::steady_simulation with the reference's residual output."""

    def __init__(self, ctx: StepContext, output_screen=10):
        self.ctx = ctx
        self.output_screen = output_screen
        self.res_scale = np.ones(4)
        self.history = []
        self.step_no = 0

    def initialize_computation(self):
        self.ctx.upload_initial()
        self.ctx.init_conservative()

    def simulate(self, total_step):
        for _ in range(total_step):
            self.step_no += 1
            self.ctx.step(1)
            if self.step_no % self.output_screen == 0:
                res = self.ctx.residual()
                if self.step_no == self.output_screen:
                    self.res_scale = np.maximum(res, 1e-20)
                r = res / self.res_scale
                self.history.append((self.step_no, float(r.max())))
        return self.history
