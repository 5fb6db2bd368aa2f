"""The workloads at their full sizes.

The oracle finishes one step of the 256x128x128 9-species block in a few
seconds on the host cores, so that block is compared with it directly (one
SSP-RK3 step at the FP64 step bar).  Beyond that the checks are properties
that do not depend on size:
* the 64-block lattice (40.4M points): per-block eager execution and the
  batched, graph-captured step give identical results bit for bit (the four
  RMS residuals after two steps, and whole blocks at the lattice's corners
  and centre), and the health check passes;
* the 384x256x256 block: the mixed-precision step with the trained surrogate
  stays within the 1e-4 bar per field of the FP64 step after 10 steps, with
  no point outside the surrogate's envelope.
"""
import os

import numpy as np
import pytest

from combustioncodegpu_amd import abi, cases
from tests.helpers import assert_close_sel

pytestmark = pytest.mark.gpu


def test_c1_full_one_step_against_oracle():
    from combustioncodegpu_amd.lib import Solver
    from oracle import pyoracle
    from oracle.pyoracle import Oracle
    used = pyoracle.set_threads(min(16, os.cpu_count() or 1))
    case = cases.c1(256, 128, 128)
    print(f"[fullsize] 256x128x128 9 species, oracle on {used} threads", flush=True)
    o = Oracle(case)
    g = Solver(case)
    for c in (o, g):
        c.upload_initial()
        c.init_conservative()
    o.step(1)
    g.step(1)
    for sel, label in ((abi.CV, "cv"), (abi.BV, "bv"), (abi.YK, "yk")):
        assert_close_sel(sel, g.download(0, sel), o.download(0, sel), 1e-12, f"256x128x128 1 step {label}")
    rc, nfail = g.check()
    assert rc == 0 and nfail == 0
    o.close()
    g.close()


def test_c3_full_modes_bit_identical():
    from combustioncodegpu_amd.lib import Solver
    case = cases.c3()
    nb = len(case.blocks)
    print(f"[fullsize] 64-block lattice: {case.n_cells} cells in {nb} blocks", flush=True)
    seq = Solver(case)
    seq.set_execution(batched=False, graph=False)
    bat = Solver(case)
    bat.set_execution(batched=True, graph=True)
    for g in (seq, bat):  # one context at a time: no kernels of two contexts overlap
        g.upload_initial()
        g.init_conservative()
        g.step(2)
        g.sync()
    rs, rb = seq.residual(), bat.residual()
    print(f"[fullsize] residuals per block {rs}, batched {rb}", flush=True)
    assert np.all(np.isfinite(rs)) and np.array_equal(rs, rb)
    for b in (0, 21, 42, nb - 1):
        for sel in (abi.CV, abi.BV, abi.YK):
            assert np.array_equal(seq.download(b, sel), bat.download(b, sel)), (b, sel)
    for g in (seq, bat):
        rc, nfail = g.check()
        assert rc == 0 and nfail == 0
        g.close()


def test_c4_full_surrogate_within_bar():
    from combustioncodegpu_amd.lib import Solver
    case = cases.c1(384, 256, 256)
    print(f"[fullsize] 384x256x256 9 species: FP64 and mixed + surrogate, 10 steps", flush=True)
    ref = Solver(case)
    case.params["precision"] = abi.MIXED
    sur = Solver(case)
    case.params["precision"] = abi.FP64
    sur.mlp_train()
    sur.set_transport(abi.TRANSPORT_MLP)
    for g in (ref, sur):
        g.set_execution(batched=True, graph=True)
        g.upload_initial()
        g.init_conservative()
        g.step(10)
        g.sync()
    br, bs = ref.download(0, abi.BV)[:, 2:-2, 2:-2, 2:-2], sur.download(0, abi.BV)[:, 2:-2, 2:-2, 2:-2]
    vmax = np.sqrt((br[1:4] ** 2).sum(axis=0)).max()
    err = {"rho": np.abs(bs[0] - br[0]).max() / np.abs(br[0]).max(),
           "velocity": np.sqrt(((bs[1:4] - br[1:4]) ** 2).sum(axis=0)).max() / vmax,
           "T": np.abs(bs[5] - br[5]).max() / np.abs(br[5]).max()}
    del br, bs
    yr, ys = ref.download(0, abi.YK)[:, 2:-2, 2:-2, 2:-2], sur.download(0, abi.YK)[:, 2:-2, 2:-2, 2:-2]
    err["Y"] = np.abs(ys - yr).max()
    stats = sur.mlp_stats()
    print(f"[fullsize] surrogate vs FP64 after 10 steps: {err}; counters {stats}", flush=True)
    for k, v in err.items():
        assert v <= 1e-4, (k, v)
    assert stats["outside"] == 0
    assert stats["checked"] > 0 and stats["check_failed"] <= stats["checked"] // 100 + 1
    for g in (ref, sur):
        rc, nfail = g.check()
        assert rc == 0 and nfail == 0
        g.close()
