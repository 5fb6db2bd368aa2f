"""Mixed precision and multiblock exchange on the GPU, against the oracle.

Mixed precision (FP32 reconstruction, AUSM+, viscous flux and transport;
FP64 Newton inversion, RK accumulation and dt/residual reductions) is
compared with the FP64 oracle per field: density, velocity (as a vector),
pressure, temperature (max relative error) and mass fractions (absolute,
they are O(1)).  The bar after 100 RK3 steps is 1e-4 for rho, |V|, T and Y.

The 2-block split of a block (inner face with shared vertex layer, on-device
exchange) is compared with the oracle's 2-block run at the FP64 step bar.
"""
import numpy as np
import pytest

from combustioncodegpu_amd import abi, cases
from tests.helpers import assert_close_sel, interior

pytestmark = pytest.mark.gpu


def field_errors(g, o, case):
    bg, bo = interior(g.download(0, abi.BV)), interior(o.download(0, abi.BV))
    vel = np.sqrt((bo[1:4] ** 2).sum(axis=0)).max()
    err = {
        "rho": float(np.abs(bg[0] - bo[0]).max() / np.abs(bo[0]).max()),
        "velocity": float(np.sqrt(((bg[1:4] - bo[1:4]) ** 2).sum(axis=0)).max() / vel),
        "p": float(np.abs(bg[4] - bo[4]).max() / np.abs(bo[4]).max()),
        "T": float(np.abs(bg[5] - bo[5]).max() / np.abs(bo[5]).max()),
    }
    if case.n_spec:
        yg, yo = interior(g.download(0, abi.YK)), interior(o.download(0, abi.YK))
        err["Y"] = float(np.abs(yg - yo).max())
    return err


def mixed_case(name):
    if name == "c1s":
        return cases.c1(24, 20, 16)
    if name == "c1s_walls":  # inflow, outflow, isothermal and adiabatic walls, slip walls, 9 species
        return cases.c1_walls(22, 18, 14)
    if name.startswith("c1s_n"):  # the other compiled species counts
        return cases.c1(16, 14, 10, n_spec=int(name[5:]))
    return cases.c0()


@pytest.mark.parametrize("name", ["c1s", "c0", "c1s_walls", "c1s_n1", "c1s_n3", "c1s_n5"])
def test_mixed_precision_per_field_error(name):
    from combustioncodegpu_amd.lib import Solver
    from oracle.pyoracle import Oracle
    case = mixed_case(name)
    mixed = mixed_case(name)
    mixed.params["precision"] = abi.MIXED
    print(f"[mixed] case={name}", flush=True)
    o, g = Oracle(case), Solver(mixed)
    for c in (o, g):
        c.upload_initial()
        c.init_conservative()
    report = {}
    done = 0
    for n in (1, 10, 100):
        o.step(n - done)
        g.step(n - done)
        done = n
        report[n] = field_errors(g, o, case)
        print(f"[mixed] {name} after {n} steps: {report[n]}", flush=True)
    rc, nfail = g.check()
    assert rc == 0 and nfail == 0
    for k in ("rho", "velocity", "T", "Y"):
        if k in report[100]:
            assert report[100][k] <= 1e-4, (k, report[100][k])
    assert report[1]["rho"] <= 1e-5


def test_two_block_exchange_parity():
    from combustioncodegpu_amd.lib import Solver
    from oracle.pyoracle import Oracle
    case = cases.split_i(cases.c1(20, 14, 12), 11)
    print("[exchange] 2-block split 11+10 of 20x14x12", flush=True)
    o, g = Oracle(case), Solver(case)
    for c in (o, g):
        c.upload_initial()
        c.init_conservative()
        c.step(3)
    for b in range(2):
        for sel in (abi.CV, abi.BV, abi.YK):
            assert_close_sel(sel, g.download(b, sel), o.download(b, sel), 3e-12, f"block {b} sel {sel}")


def test_two_block_equals_one_block_on_gpu():
    from combustioncodegpu_amd.lib import Solver
    one = cases.c1(20, 14, 12)
    two = cases.split_i(one, 11)
    g1, g2 = Solver(one), Solver(two)
    for c in (g1, g2):
        c.upload_initial()
        c.init_conservative()
        c.step(3)
    a = g1.download(0, abi.CV)
    A, B = g2.download(0, abi.CV), g2.download(1, abi.CV)
    ia, ib = a[..., 2:-2, 2:-2, 2:2 + 11], a[..., 2:-2, 2:-2, 2 + 10:-2]
    assert_close_sel(abi.CV, A[..., 2:-2, 2:-2, 2:-2], ia, 1e-13, "block A")
    assert_close_sel(abi.CV, B[..., 2:-2, 2:-2, 2:-2], ib, 1e-13, "block B")
