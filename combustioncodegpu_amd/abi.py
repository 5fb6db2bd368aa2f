"""ctypes mirror of include/cfx.h and marshalling of a :class:`Case` into it.

The structs follow the C header field by field; `marshal_case` builds them
from numpy arrays and keeps those arrays alive for as long as the returned
holder lives (the C side copies what it needs during ``cfx_create``).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

CFX_MAX_SPEC = 16

CFX_E_ARG = -1
CFX_E_NEWTON = -2
CFX_E_NONFINITE = -3
CFX_E_SURROGATE = -4
CFX_E_UNSUPPORTED = -5

BC_WALL, BC_SYMMETRY, BC_INFLOW, BC_OUTFLOW = 2, 3, 5, 6

Q_METRIC_AVG = 0x1
Q_HV_RHOD_J = 0x2
Q_NEWTON_REL = 0x4
QUIRKS_REFERENCE = Q_METRIC_AVG | Q_HV_RHOD_J

FP64, MIXED = 0, 1
TRANSPORT_WILKE, TRANSPORT_MLP = 0, 1

CV, BV, YK, PROPS, DQ, METRIC, DTLOCAL, COORD = range(8)

_dp = C.POINTER(C.c_double)


class CfxBC(C.Structure):
    _fields_ = [("range_start", C.c_int * 3), ("range_end", C.c_int * 3), ("type", C.c_int), ("cond", C.c_int)]


class CfxInner(C.Structure):
    _fields_ = [("src", C.c_int * 6), ("tar", C.c_int * 6), ("target_block", C.c_int)]


class CfxBlock(C.Structure):
    _fields_ = [
        ("mx", C.c_int), ("my", C.c_int), ("mz", C.c_int),
        ("x", _dp), ("y", _dp), ("z", _dp),
        ("n_bc", C.c_int), ("bc", C.POINTER(CfxBC)),
        ("n_inner", C.c_int), ("inner", C.POINTER(CfxInner)),
    ]


class CfxInflow(C.Structure):
    _fields_ = [
        ("mach", C.c_double), ("pressure", C.c_double), ("temperature", C.c_double),
        ("velocity", C.c_double), ("density", C.c_double),
        ("u", C.c_double), ("v", C.c_double), ("w", C.c_double),
        ("yk", _dp),
    ]


class CfxWall(C.Structure):
    _fields_ = [("adiabatic", C.c_int), ("temperature", C.c_double)]


class CfxParams(C.Structure):
    _fields_ = [
        ("inviscid_scheme", C.c_int), ("reconstruction", C.c_int), ("limiter", C.c_int),
        ("viscous_order", C.c_int), ("temporal_scheme", C.c_int), ("rk_stages", C.c_int),
        ("cfl", C.c_double), ("prandtl", C.c_double), ("schmidt", C.c_double),
        ("gamma_air", C.c_double), ("r_air", C.c_double),
        ("precision", C.c_int), ("transport", C.c_int), ("quirks", C.c_int),
        ("n_inflow", C.c_int), ("inflow", C.POINTER(CfxInflow)),
        ("n_wall", C.c_int), ("wall", C.POINTER(CfxWall)),
        ("lusgs", C.c_int), ("lusgs_omega", C.c_double),
    ]


class CfxSpecies(C.Structure):
    _fields_ = [
        ("n_spec", C.c_int),
        ("mw", _dp), ("t_low", _dp), ("t_mid", _dp), ("t_high", _dp),
        ("coeff_lo", _dp), ("coeff_hi", _dp), ("eps_k", _dp), ("sigma", _dp),
    ]


class CfxMlpConfig(C.Structure):
    _fields_ = [
        ("n_samples", C.c_longlong), ("holdout", C.c_double), ("epochs", C.c_int), ("batch", C.c_int),
        ("lr", C.c_double), ("lr_final", C.c_double), ("t_min", C.c_double), ("t_max", C.c_double),
        ("p_min", C.c_double), ("p_max", C.c_double), ("dirichlet_alpha", C.c_double), ("check_tol", C.c_double),
        ("check_every", C.c_int), ("seed", C.c_ulonglong), ("envelope_quantile", C.c_double),
    ]


class CfxMlpReport(C.Structure):
    _fields_ = [
        ("samples_ms", C.c_double), ("train_ms", C.c_double), ("epochs", C.c_int), ("steps", C.c_int),
        ("loss_first", C.c_double), ("loss_last", C.c_double), ("val_rmse", C.c_double),
        ("val_max_rel", C.c_double * 3), ("val_p99_rel", C.c_double * 3), ("n_train", C.c_longlong),
        ("n_val", C.c_longlong), ("n_params", C.c_int), ("val_answered", C.c_double),
    ]


def dptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"], "expected contiguous float64"
    return a.ctypes.data_as(_dp)


class Marshalled:
    """Holds the ctypes structs of a case and every array they point into."""

    def __init__(self, case):
        self._keep = []
        p = case.params
        inflows = case.inflow
        self.inflow_arr = (CfxInflow * max(1, len(inflows)))()
        for i, f in enumerate(inflows):
            r = self.inflow_arr[i]
            for key in ("mach", "pressure", "temperature", "velocity", "density", "u", "v", "w"):
                setattr(r, key, float(f.get(key, {"mach": -1, "pressure": 101325, "temperature": -1,
                                                   "velocity": 0, "density": -1, "u": 1, "v": 0, "w": 0}[key])))
            if f.get("yk") is not None:
                y = np.ascontiguousarray(f["yk"], dtype=np.float64)
                self._keep.append(y)
                r.yk = dptr(y)
        walls = case.wall
        self.wall_arr = (CfxWall * max(1, len(walls)))()
        for i, w in enumerate(walls):
            self.wall_arr[i].adiabatic = int(w.get("adiabatic", 0))
            self.wall_arr[i].temperature = float(w.get("temperature", 300.0))
        self.params = CfxParams(
            inviscid_scheme=p.get("inviscid_scheme", 3), reconstruction=p.get("reconstruction", 2),
            limiter=p.get("limiter", 0), viscous_order=p.get("viscous_order", 2),
            temporal_scheme=p.get("temporal_scheme", 0), rk_stages=p.get("rk_stages", 3),
            cfl=p.get("cfl", 0.5), prandtl=p.get("prandtl", 0.72), schmidt=p.get("schmidt", 0.9),
            gamma_air=p.get("gamma_air", 1.4), r_air=p.get("r_air", 8314.462618 / 28.9647),
            precision=p.get("precision", FP64), transport=p.get("transport", TRANSPORT_WILKE),
            quirks=p.get("quirks", QUIRKS_REFERENCE),
            n_inflow=len(inflows), inflow=self.inflow_arr,
            n_wall=len(walls), wall=self.wall_arr,
            lusgs=p.get("lusgs", 0), lusgs_omega=p.get("lusgs_omega", 0.0),
        )
        sp = case.species
        self.species = CfxSpecies(n_spec=0)
        if sp is not None:
            arrs = {}
            for key in ("mw", "t_low", "t_mid", "t_high", "coeff_lo", "coeff_hi", "eps_k", "sigma"):
                a = np.ascontiguousarray(sp[key], dtype=np.float64).ravel()
                self._keep.append(a)
                arrs[key] = dptr(a)
            self.species = CfxSpecies(n_spec=len(sp["mw"]), **arrs)
        self.blocks = (CfxBlock * len(case.blocks))()
        for b, blk in enumerate(case.blocks):
            xyz = []
            for a in (blk.x, blk.y, blk.z):
                a = np.ascontiguousarray(a, dtype=np.float64).ravel()
                self._keep.append(a)
                xyz.append(a)
            bcs = (CfxBC * max(1, len(blk.bc)))()
            for f, (rs, re, typ, cond) in enumerate(blk.bc):
                bcs[f].range_start[:] = list(rs)
                bcs[f].range_end[:] = list(re)
                bcs[f].type = typ
                bcs[f].cond = cond
            inn = (CfxInner * max(1, len(blk.inner)))()
            for f, (src, tar, tb) in enumerate(blk.inner):
                inn[f].src[:] = list(src)
                inn[f].tar[:] = list(tar)
                inn[f].target_block = tb
            self._keep += [bcs, inn]
            self.blocks[b] = CfxBlock(mx=blk.mx, my=blk.my, mz=blk.mz, x=dptr(xyz[0]), y=dptr(xyz[1]),
                                      z=dptr(xyz[2]), n_bc=len(blk.bc), bc=bcs, n_inner=len(blk.inner), inner=inn)
        self.n_block = len(case.blocks)


def declare_api(lib, prefix: str):
    """Declare argtypes/restype of the step API under a prefix (cfx_ / orc_)."""
    vp = C.c_void_p
    sig = {
        "create": ([C.POINTER(CfxParams), C.POINTER(CfxSpecies), C.POINTER(CfxBlock), C.c_int, C.POINTER(vp)], C.c_int),
        "destroy": ([vp], None),
        "last_error": ([vp], C.c_char_p),
        "upload_primitive": ([vp, C.c_int, _dp, _dp], C.c_int),
        "upload_state": ([vp, C.c_int, _dp, _dp, _dp], C.c_int),
        "download": ([vp, C.c_int, C.c_int, _dp], C.c_int),
        "init_conservative": ([vp], C.c_int),
        "store_last_step": ([vp], C.c_int),
        "rhs": ([vp, C.c_int], C.c_int),
        "time_step": ([vp, C.c_int, _dp], C.c_int),
        "update": ([vp, C.c_int], C.c_int),
        "lusgs": ([vp], C.c_int),
        "apply_bc": ([vp], C.c_int),
        "exchange": ([vp], C.c_int),
        "eliminate_k_gradient": ([vp], C.c_int),
        "update_properties": ([vp], C.c_int),
        "residual": ([vp, _dp], C.c_int),
        "step": ([vp, C.c_int], C.c_int),
        "get_dt": ([vp, _dp], C.c_int),
        "set_dt": ([vp, C.c_double], C.c_int),
    }
    for name, (args, res) in sig.items():
        fn = getattr(lib, prefix + name)
        fn.argtypes = args
        fn.restype = res
