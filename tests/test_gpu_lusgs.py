"""The implicit LU-SGS step on the GPU against the oracle.

The reference declares LU-SGS (TemporalScheme.cuh:19-21) without a body; the
oracle restates the published matrix-free scheme and is pinned in
tests/test_oracle_lusgs.py.  Here the HIP path (coefficient kernel and the
forward / backward hyperplane sweeps) is compared with it:
  - the increment dQ of one solve, from the same RHS and local time step;
  - whole steps (RHS, local dt, LU-SGS, update, ghost fill, properties) on
    a 9-species 3-D block, the 2-D air case, 2 species and a 12-block
    lattice, at the FP64 step bar (1e-12 per step, normwise per quantity);
  - per-block, batched and graph-replayed execution give identical bytes.
"""
import numpy as np
import pytest

from combustioncodegpu_amd import abi, cases
from tests.helpers import assert_close_sel
from tests.test_gpu_parity import compare_state

pytestmark = pytest.mark.gpu

TOL_STEP = 1e-12


def lusgs(case, cfl=5.0, omega=1.0):
    case.params["rk_stages"] = 1
    case.params["lusgs"] = 1
    case.params["lusgs_omega"] = omega
    case.params["cfl"] = cfl
    return case


def make(name):
    if name == "c1s":
        return lusgs(cases.c1(18, 14, 12, rk_stages=1))
    if name == "c1s_curved_ns2":
        return lusgs(cases.c1(16, 14, 10, n_spec=2, perturb=0.3, rk_stages=1), cfl=10.0, omega=1.5)
    if name == "c0":
        return lusgs(cases.c0(rk_stages=1))
    if name == "c1s_walls":
        return lusgs(cases.c1_walls(16, 14, 12, rk_stages=1), cfl=3.0)
    raise KeyError(name)


def pair(case):
    from combustioncodegpu_amd.lib import Solver
    from oracle.pyoracle import Oracle
    o, g = Oracle(case), Solver(case)
    for c in (o, g):
        c.upload_initial()
        c.init_conservative()
    return o, g


@pytest.mark.parametrize("name", ["c1s", "c0", "c1s_walls"])
def test_lusgs_increment_parity(name):
    case = make(name)
    print(f"[lusgs] increment case={name}", flush=True)
    o, g = pair(case)
    for c in (o, g):
        c.rhs()
        c.time_step(False)
        c.lusgs()
    for b in range(len(case.blocks)):
        assert_close_sel(abi.DQ, g.download(b, abi.DQ), o.download(b, abi.DQ), TOL_STEP, f"dQ blk {b}")


@pytest.mark.parametrize("name,steps", [("c1s", 3), ("c1s_curved_ns2", 3), ("c0", 5), ("c1s_walls", 3)])
def test_lusgs_step_parity(name, steps):
    case = make(name)
    print(f"[lusgs] steps case={name} n={steps}", flush=True)
    o, g = pair(case)
    for s in range(1, steps + 1):
        o.step(1)
        g.step(1)
        compare_state(case, o, g, TOL_STEP * s, f"LU-SGS step {s}")
    rc, nfail = g.check()
    assert rc == 0 and nfail == 0


def test_lusgs_lattice_modes():
    """A 12-block lattice: parity with the oracle after 2 steps, and
    per-block, batched and graph-replayed runs identical."""
    from combustioncodegpu_amd.lib import Solver
    from oracle.pyoracle import Oracle
    case = lusgs(cases.lattice([[9, 14, 11], [10, 8], [7, 12]], rk_stages=1))
    nb = len(case.blocks)
    print(f"[lusgs] lattice blocks={nb}", flush=True)
    o = Oracle(case)
    o.upload_initial()
    o.init_conservative()
    o.step(2)
    runs = {}
    for mode, (batched, graph) in {"seq": (False, False), "bat": (True, False), "graph": (True, True)}.items():
        g = Solver(case)
        g.set_execution(batched=batched, graph=graph)
        g.upload_initial()
        g.init_conservative()
        g.step(2)
        runs[mode] = [np.concatenate([g.download(b, s).ravel() for s in (abi.CV, abi.BV, abi.YK)])
                      for b in range(nb)]
        if mode == "bat":
            for b in range(nb):
                for sel in (abi.CV, abi.BV, abi.YK):
                    assert_close_sel(sel, g.download(b, sel), o.download(b, sel), 2 * TOL_STEP, f"blk {b} {sel}")
        g.close()
    for b in range(nb):
        assert np.array_equal(runs["seq"][b], runs["bat"][b])
        assert np.array_equal(runs["bat"][b], runs["graph"][b])


def test_lusgs_rejects_unsupported_forms():
    from combustioncodegpu_amd.driver import CfxError
    from combustioncodegpu_amd.lib import Solver
    c = lusgs(cases.c1(10, 8, 6, rk_stages=1))
    c.params["rk_stages"] = 3
    with pytest.raises(CfxError):
        Solver(c)
    c = lusgs(cases.c1(10, 8, 6, rk_stages=1))
    c.params["precision"] = abi.MIXED
    with pytest.raises(CfxError):
        Solver(c)
