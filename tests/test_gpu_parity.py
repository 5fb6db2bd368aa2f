"""Parity of the HIP step (through the C ABI) with the CPU oracle.

Every comparison is normwise: max|gpu - oracle| over the compared range
divided by the quantity's scale (tests/helpers.py: momentum as a vector,
mass fractions by 1, every other quantity by its own maximum).  The bar is the target tolerance of 1e-12 per
step on conserved variables (FP64 path); multi-step runs allow the per-step
bar times the step count.  Inputs are the seeded workloads of
combustioncodegpu_amd.cases at sizes the oracle finishes in seconds.
"""
import numpy as np
import pytest

from combustioncodegpu_amd import abi, cases
from tests.helpers import assert_close, assert_close_sel, interior

pytestmark = pytest.mark.gpu

TOL_STEP = 1e-12


def make(name):
    if name == "c0":
        return cases.c0()
    if name == "c0_euler":
        return cases.c0(rk_stages=1)
    if name == "c0_va":
        return cases.c0(limiter=1, mx=64, my=32)
    if name == "c1s":
        return cases.c1(24, 20, 16)
    if name == "c1s_minmod":
        return cases.c1(20, 16, 12, limiter=0)
    if name == "c1s_ns2":
        return cases.c1(20, 18, 14, n_spec=2)
    if name == "c1s_air":
        return cases.c1(20, 18, 14, n_spec=0)
    if name == "c1s_curv_ref":
        return cases.c1(20, 16, 12, perturb=0.3)
    if name == "c1s_curv_corr":
        return cases.c1(20, 16, 12, perturb=0.3, quirks=0)
    if name == "c1s_euler":
        return cases.c1(18, 16, 12, rk_stages=1)
    if name == "c1s_inviscid":
        return cases.c1(18, 16, 12, viscous=0)
    if name == "c1s_walls":
        return cases.c1_walls(22, 18, 14)
    if name == "c1s_first":
        return cases.c1(18, 14, 12, reconstruction=1)
    if name.startswith("c1s_n") and name[5:].isdigit():  # the other compiled species counts
        return cases.c1(16, 14, 10, n_spec=int(name[5:]))
    raise KeyError(name)


CASES = ["c0", "c0_euler", "c0_va", "c1s", "c1s_minmod", "c1s_ns2", "c1s_air", "c1s_curv_ref",
         "c1s_curv_corr", "c1s_euler", "c1s_inviscid", "c1s_walls", "c1s_first", "c1s_n1", "c1s_n3", "c1s_n5"]


def pair(name):
    from combustioncodegpu_amd.lib import Solver
    from oracle.pyoracle import Oracle
    case = make(name)
    print(f"[parity] case={name} blocks={[(b.mx, b.my, b.mz) for b in case.blocks]} ns={case.n_spec}", flush=True)
    o = Oracle(case)
    g = Solver(case)
    for c in (o, g):
        c.upload_initial()
        c.init_conservative()
    return case, o, g


def compare_state(case, o, g, tol, what):
    dim = 2 if case.blocks[0].mz == 1 else 3
    for b in range(len(case.blocks)):
        for sel, label in ((abi.CV, "cv"), (abi.BV, "bv"), (abi.YK, "yk"), (abi.PROPS, "props")):
            if sel == abi.YK and case.n_spec == 0:
                continue
            ga, oa = g.download(b, sel), o.download(b, sel)
            assert_close_sel(sel, interior(ga, dim=dim), interior(oa, dim=dim), tol, f"{what} {label} interior blk {b}")
            assert_close_sel(sel, ga, oa, tol, f"{what} {label} padded blk {b}")


@pytest.mark.parametrize("name", CASES)
def test_init_parity(name):
    case, o, g = pair(name)
    for b in range(len(case.blocks)):
        assert_close(g.download(b, abi.METRIC), o.download(b, abi.METRIC), 1e-14, "metric")
    compare_state(case, o, g, 1e-13, "init")


@pytest.mark.parametrize("name", CASES)
def test_rhs_and_dt_parity(name):
    case, o, g = pair(name)
    for c in (o, g):
        c.rhs(-1)
    for b in range(len(case.blocks)):
        assert_close_sel(abi.DQ, g.download(b, abi.DQ), o.download(b, abi.DQ), 1e-12, "dq")
    dto = o.time_step(True)
    dtg = g.time_step(True)
    assert abs(dtg - dto) <= 1e-13 * dto
    for b in range(len(case.blocks)):
        assert_close(g.download(b, abi.DTLOCAL), o.download(b, abi.DTLOCAL), 1e-13, "dt_local")


@pytest.mark.parametrize("name", CASES)
def test_one_step_parity(name):
    case, o, g = pair(name)
    o.step(1)
    g.step(1)
    compare_state(case, o, g, TOL_STEP, "1 step")
    rc, nfail = g.check()
    assert rc == 0 and nfail == 0


@pytest.mark.parametrize("name,steps", [("c0", 20), ("c1s", 8), ("c1s_curv_corr", 5), ("c0_euler", 20),
                                        ("c1s_walls", 5)])
def test_multi_step_parity(name, steps):
    case, o, g = pair(name)
    o.step(steps)
    g.step(steps)
    compare_state(case, o, g, TOL_STEP * steps, f"{steps} steps")
    # residuals are RMS of step differences (d rho, d|V|, d p, d T), i.e.
    # cancellations of O(q) values: compare against the state bar times q
    ro, rg = o.residual(), g.residual()
    bv = o.download(0, abi.BV)
    q = np.array([np.abs(bv[0]).max(), np.sqrt((bv[1:4] ** 2).sum(axis=0)).max(), np.abs(bv[4]).max(),
                  np.abs(bv[5]).max()])
    assert np.all(np.abs(rg - ro) <= TOL_STEP * steps * q), (rg, ro)
    assert abs(g.get_dt() - o.get_dt()) <= 1e-12 * o.get_dt()


def test_stage_pieces_parity():
    """The step pieces called one by one (the reference Driver loop shape)."""
    case, o, g = pair("c1s")
    for c in (o, g):
        c.store_last_step()
        for s in range(3):
            c.rhs(-1)
            if s == 0:
                c.time_step(True)
            c.update(s)
            c.apply_bc()
            c.exchange()
            c.eliminate_k_gradient()
            c.update_properties()
    compare_state(case, o, g, TOL_STEP, "pieces")


def folded_ghost_case():
    """A block whose j-lines spread with i (y = j dy (1 + 0.6 i)): the linear
    ghost extrapolation folds at i = -2 (J <= 0 there), where the reference
    copies the metrics of i = -1."""
    c = cases.c1(12, 10, 8)
    k, j, i = np.meshgrid(np.arange(8), np.arange(10), np.arange(12), indexing="ij")
    c.blocks[0].y = np.ascontiguousarray((j * 1e-3 * (1 + 0.6 * i)).ravel(), dtype=np.float64)
    return c


def test_metric_fallback_on_folded_ghosts():
    """The device metric pass finds the non-positive Jacobians and the host
    recomputes the block with the fallback: metrics as the oracle's."""
    from combustioncodegpu_amd.lib import Solver
    from oracle.pyoracle import Oracle
    case = folded_ghost_case()
    print("[parity] folded ghost grid 12x10x8", flush=True)
    mo = Oracle(case).download(0, abi.METRIC)
    g = Solver(case)
    mg = g.download(0, abi.METRIC)
    g.close()
    assert np.array_equal(mg[:, 2:-2, 3:-3, 0], mg[:, 2:-2, 3:-3, 1])   # i = -2 took i = -1's values
    assert_close(mg, mo, 1e-14, "metrics with fallback")
