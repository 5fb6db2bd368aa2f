"""TEST INFRASTRUCTURE: ctypes binding of the CPU oracle (oracle/_build).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module; it is the checker, never the product path.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

from combustioncodegpu_amd import abi
from combustioncodegpu_amd.driver import StepContext

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")
_lib = None


def build(force=False):
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-C", HERE, "-s"], check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = C.CDLL(LIB_PATH)
        abi.declare_api(_lib, "orc_")
        _lib.orc_set_threads.argtypes = [C.c_int]
        _lib.orc_set_threads.restype = C.c_int
        _lib.orc_set_dq.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_double)]
        _lib.orc_set_dq.restype = C.c_int
        _lib.orc_newton_failures.argtypes = [C.c_void_p]
        _lib.orc_newton_failures.restype = C.c_longlong
        dp = C.POINTER(C.c_double)
        _lib.orc_enthalpy_cp.argtypes = [C.POINTER(abi.CfxSpecies), C.c_double, dp, dp]
        _lib.orc_transport.argtypes = [C.POINTER(abi.CfxSpecies), C.c_double, C.c_double, dp, dp, dp, dp]
        _lib.orc_newton_temperature.argtypes = [C.POINTER(abi.CfxSpecies), C.c_int, dp, C.c_double, C.c_double, dp]
        _lib.orc_newton_temperature.restype = C.c_int
        _lib.orc_limiter.argtypes = [C.c_int, C.c_double, C.c_double]
        _lib.orc_limiter.restype = C.c_double
        _lib.orc_sutherland.argtypes = [C.c_double]
        _lib.orc_sutherland.restype = C.c_double
        _lib.orc_inviscid_face.argtypes = [C.POINTER(abi.CfxSpecies), C.POINTER(abi.CfxParams), dp, dp, dp, dp]
        _lib.orc_face_flux.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, dp]
        _lib.orc_face_flux.restype = C.c_int
        fp = C.POINTER(C.c_float)
        _lib.orc_mlp_forward.argtypes = [fp, C.c_int, C.c_int, fp, C.c_int, C.c_int, dp]
        _lib.orc_mlp_forward.restype = C.c_int
        _lib.orc_mlp_loss_grad.argtypes = [fp, C.c_int, C.c_int, fp, fp, C.c_int, C.c_int, dp, dp]
        _lib.orc_mlp_loss_grad.restype = C.c_int
    return _lib


def set_threads(n):
    return lib().orc_set_threads(int(n))


class Oracle(StepContext):
    def __init__(self, case):
        super().__init__(lib(), "orc_", case)

    def newton_failures(self):
        return self.lib.orc_newton_failures(self.h)

    def set_dq(self, blk, dq):
        a = np.ascontiguousarray(dq, dtype=np.float64)
        rc = self.lib.orc_set_dq(self.h, blk, abi.dptr(a))
        if rc:
            raise RuntimeError(f"orc_set_dq: {rc}")

    def face_flux(self, blk, d, i, j, k, viscous):
        """Face flux between interior point (i,j,k) and (i,j,k)+e_d."""
        out = np.zeros(5 + self.case.n_spec)
        rc = self.lib.orc_face_flux(self.h, blk, d, i, j, k, 1 if viscous else 0, abi.dptr(out))
        if rc:
            raise RuntimeError(f"orc_face_flux: {rc}")
        return out


class SpeciesHolder:
    """Marshalled species struct for the point-wise oracle functions."""

    def __init__(self, sp):
        self._keep = []
        arrs = {}
        for key in ("mw", "t_low", "t_mid", "t_high", "coeff_lo", "coeff_hi", "eps_k", "sigma"):
            a = np.ascontiguousarray(sp[key], dtype=np.float64).ravel()
            self._keep.append(a)
            arrs[key] = abi.dptr(a)
        self.n = len(sp["mw"])
        self.s = abi.CfxSpecies(n_spec=self.n, **arrs)


def enthalpy_cp(sp, t):
    h_ = SpeciesHolder(sp)
    h = np.zeros(h_.n)
    cp = np.zeros(h_.n)
    lib().orc_enthalpy_cp(C.byref(h_.s), float(t), abi.dptr(h), abi.dptr(cp))
    return h, cp


def transport(sp, t, Y, schmidt=0.9):
    h_ = SpeciesHolder(sp)
    Y = np.ascontiguousarray(Y, dtype=np.float64)
    mu = np.zeros(1)
    kap = np.zeros(1)
    rd = np.zeros(h_.n)
    lib().orc_transport(C.byref(h_.s), schmidt, float(t), abi.dptr(Y), abi.dptr(mu), abi.dptr(kap), abi.dptr(rd))
    return mu[0], kap[0], rd


def newton_temperature(sp, Y, e, t_guess, quirks=0):
    h_ = SpeciesHolder(sp)
    Y = np.ascontiguousarray(Y, dtype=np.float64)
    t = np.zeros(1)
    it = lib().orc_newton_temperature(C.byref(h_.s), quirks, abi.dptr(Y), float(e), float(t_guess), abi.dptr(t))
    return t[0], it


def limiter(tag, a, b):
    return lib().orc_limiter(tag, a, b)


def sutherland(t):
    return lib().orc_sutherland(t)


def inviscid_face(sp, params, pv, metric, jac):
    """pv: (4, 5+ns) primitive points; metric: (2,3); jac: (2,)."""
    ns = 0 if sp is None else len(sp["mw"])
    h_ = SpeciesHolder(sp) if sp is not None else None
    s = h_.s if h_ is not None else abi.CfxSpecies(n_spec=0)
    prm = abi.CfxParams(reconstruction=params.get("reconstruction", 2), limiter=params.get("limiter", 0),
                        gamma_air=params.get("gamma_air", 1.4), r_air=params.get("r_air", 287.0))
    pv = np.ascontiguousarray(pv, dtype=np.float64)
    metric = np.ascontiguousarray(metric, dtype=np.float64)
    jac = np.ascontiguousarray(jac, dtype=np.float64)
    f = np.zeros(5 + ns)
    lib().orc_inviscid_face(C.byref(s), C.byref(prm), abi.dptr(pv), abi.dptr(metric), abi.dptr(jac), abi.dptr(f))
    return f


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def mlp_forward(params, nin, nout, x, emulate=True):
    """Network outputs y[n][16] of inputs x[n][16] (float rows).  emulate:
    True (or 1) the GPU's inference rounding, False (0) exact, 2 the GPU's
    folded form of the network without rounding."""
    p = np.ascontiguousarray(params, dtype=np.float32)
    xx = np.ascontiguousarray(x, dtype=np.float32)
    y = np.zeros((xx.shape[0], 16))
    rc = lib().orc_mlp_forward(_fp(p), nin, nout, _fp(xx), xx.shape[0], int(emulate), abi.dptr(y))
    assert rc == 0
    return y


def mlp_loss_grad(params, nin, nout, x, t, emulate=True):
    """(loss, gradient) of sum (y - t)^2 / (n nout)."""
    p = np.ascontiguousarray(params, dtype=np.float32)
    xx = np.ascontiguousarray(x, dtype=np.float32)
    tt = np.ascontiguousarray(t, dtype=np.float32)
    g = np.zeros(p.size)
    loss = C.c_double(0.0)
    rc = lib().orc_mlp_loss_grad(_fp(p), nin, nout, _fp(xx), _fp(tt), xx.shape[0], 1 if emulate else 0,
                                 abi.dptr(g), C.byref(loss))
    assert rc == 0
    return loss.value, g
