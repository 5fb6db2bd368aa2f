"""Batched multiblock execution on the GPU.

Lattices of uneven blocks joined by inner faces (shared vertex layers,
on-device exchange) run through the batched path: one launch per kernel
over all blocks, boundary and exchange face operations grouped into levels
of operations that touch disjoint points.  Checked here:
  - parity with the oracle's multiblock step (one op at a time, in the
    reference's order) at the FP64 step bar, for FP64;
  - per-block, batched and graph-replayed execution give identical bytes
    (the schedule keeps the order of every overlapping pair of operations,
    so batching may not change a single bit), graph replays interleaved
    with eager work and checked after each;
  - every RHS tile variant and persistence factor, and the autotuner's
    choice, give identical bytes (FP64 and mixed precision);
  - the mixed-precision path on a lattice stays within its error bar.
"""
import numpy as np
import pytest

from combustioncodegpu_amd import abi, cases
from tests.helpers import assert_close_sel

pytestmark = pytest.mark.gpu

TOL_STEP = 1e-12

SMALL = [[9, 14, 11], [10, 8], [7, 12]]   # 3 x 2 x 2 = 12 blocks


def run(case, steps, batched=True, graph=False, setup=None):
    from combustioncodegpu_amd.lib import Solver
    g = Solver(case)
    g.set_execution(batched=batched, graph=graph)
    if setup:
        setup(g)
    g.upload_initial()
    g.init_conservative()
    g.step(steps)
    return g


def state(g, n_blocks):
    return [np.concatenate([g.download(b, abi.CV).ravel(), g.download(b, abi.BV).ravel(),
                            g.download(b, abi.YK).ravel()]) for b in range(n_blocks)]


@pytest.mark.parametrize("n_spec", [9, 2])
def test_lattice_matches_oracle(n_spec):
    from oracle.pyoracle import Oracle
    case = cases.lattice(SMALL, n_spec=n_spec)
    nb = len(case.blocks)
    print(f"[batched] lattice {SMALL} ns={n_spec} blocks={nb}", flush=True)
    o = Oracle(case)
    o.upload_initial()
    o.init_conservative()
    g = run(case, 0)
    n_ops, n_lev = g.schedule_info()
    print(f"[batched] ghost-fill ops {n_ops} in {n_lev} levels", flush=True)
    assert n_lev < n_ops
    for steps in (1, 2):
        o.step(1)
        g.step(1)
        for b in range(nb):
            for sel in (abi.CV, abi.BV, abi.YK):
                if sel == abi.YK and n_spec == 0:
                    continue
                assert_close_sel(sel, g.download(b, sel), o.download(b, sel), TOL_STEP * steps,
                                 f"block {b} sel {sel} after {steps} steps")
    rc, nfail = g.check()
    assert rc == 0 and nfail == 0


def test_c3_lattice_reduced_extents_against_oracle():
    """The 64-block C3 lattice (4 x 4 x 4 blocks, seeded face-matched graph)
    at reduced extents drawn from [16, 24]: batched and graph-captured on the
    GPU against the oracle's one-op-at-a-time multiblock step, 2 steps at
    the FP64 step bar."""
    from oracle.pyoracle import Oracle
    case = cases.c3(lo=16, hi=24)
    nb = len(case.blocks)
    assert nb == 64
    print(f"[batched] c3 lattice at extents [16, 24]: {nb} blocks, {case.n_cells} cells", flush=True)
    o = Oracle(case)
    o.upload_initial()
    o.init_conservative()
    o.step(2)
    g = run(case, 2, batched=True, graph=True)
    for b in range(nb):
        for sel in (abi.CV, abi.BV, abi.YK):
            assert_close_sel(sel, g.download(b, sel), o.download(b, sel), TOL_STEP * 2, f"block {b} sel {sel}")
    rc, nfail = g.check()
    assert rc == 0 and nfail == 0


def test_execution_modes_identical():
    case = cases.lattice(SMALL, n_spec=9)
    nb = len(case.blocks)
    seq = run(case, 0, batched=False)
    bat = run(case, 0, batched=True)
    gra = run(case, 0, batched=True, graph=True)
    gseq = run(case, 0, batched=False, graph=True)
    for rep in range(3):
        print(f"[batched] modes, replay {rep}", flush=True)
        for g in (seq, bat, gra, gseq):
            g.step(2)
        ref = state(seq, nb)
        for g, name in ((bat, "batched"), (gra, "batched graph"), (gseq, "per-block graph")):
            got = state(g, nb)
            for b in range(nb):
                assert np.array_equal(got[b], ref[b]), f"{name}: block {b} differs after replay {rep}"
        # eager work on the graph contexts between replays
        for g in (gra, gseq):
            g.update_properties()
            g.sync()
    res = [g.residual() for g in (seq, bat, gra)]
    for r in res[1:]:
        assert np.allclose(r, res[0], rtol=1e-13, atol=0)


def test_tile_variants_and_autotune_identical():
    case = cases.lattice(SMALL, n_spec=9)
    nb = len(case.blocks)
    ref = state(run(case, 2), nb)
    from combustioncodegpu_amd.lib import Solver
    probe = Solver(case)
    nv = probe.rhs_variants()
    probe.close()
    for v in range(nv):
        for p in (0, 1):
            print(f"[batched] tiles variant {v} persist {p}", flush=True)
            g = run(case, 2, setup=lambda s: s.set_rhs_tiles([v] * 3, [p] * 3))
            got = state(g, nb)
            g.close()
            for b in range(nb):
                assert np.array_equal(got[b], ref[b]), f"variant {v} persist {p} block {b}"
    g = run(case, 0)
    var, per, ms = g.autotune(reps=2)
    print(f"[batched] autotune choice {var} {per} {ms}", flush=True)
    g.upload_initial()
    g.init_conservative()
    g.step(2)
    got = state(g, nb)
    for b in range(nb):
        assert np.array_equal(got[b], ref[b]), f"autotuned block {b}"


def test_mixed_tile_variants_agree():
    """Mixed precision: every RHS variant gives the state of variant 0 after
    two steps, on the lattice and on a block whose extents are not multiples
    of any tile.  The pencil tiles evaluate the same expressions and agree
    bit for bit; the bar is set at 1e-6 of each variable's maximum (the size
    of the mixed path's own FP32 rounding after one step, 2.6e-7 in velocity
    against FP64) so that a variant compiled with another multiply-add fusion
    would still pass while a wrong index or stencil point (order one) fails."""
    def fields(g, nb):
        return [[g.download(b, sel) for sel in (abi.CV, abi.BV, abi.YK)] for b in range(nb)]

    for case in (cases.lattice(SMALL, n_spec=9), cases.c1(70, 37, 41)):
        case.params["precision"] = abi.MIXED
        nb = len(case.blocks)
        g = run(case, 2)
        ref = fields(g, nb)
        nv = g.rhs_variants()
        g.close()
        for v in range(1, nv):
            g = run(case, 2, setup=lambda s: s.set_rhs_tiles([v] * 3, [0] * 3))
            got = fields(g, nb)
            g.close()
            worst = 0.0
            for b in range(nb):
                for a, r in zip(got[b], ref[b]):
                    for x, y in zip(a, r):  # per variable
                        worst = max(worst, float(np.abs(x - y).max() / max(np.abs(y).max(), 1e-300)))
            print(f"[batched] mixed tiles variant {v} on {nb} blocks: max relative difference {worst:.2e}",
                  flush=True)
            assert worst <= 1e-6, f"mixed variant {v}: {worst}"


def test_lattice_mixed_precision_error():
    from oracle.pyoracle import Oracle
    case = cases.lattice(SMALL, n_spec=9)
    mixed = cases.lattice(SMALL, n_spec=9)
    mixed.params["precision"] = abi.MIXED
    o = Oracle(case)
    o.upload_initial()
    o.init_conservative()
    o.step(10)
    g = run(mixed, 10, graph=True)
    worst = {}
    for b in range(len(case.blocks)):
        bg, bo = g.download(b, abi.BV), o.download(b, abi.BV)
        yg, yo = g.download(b, abi.YK), o.download(b, abi.YK)
        worst["rho"] = max(worst.get("rho", 0), float(np.abs(bg[0] - bo[0]).max() / np.abs(bo[0]).max()))
        worst["T"] = max(worst.get("T", 0), float(np.abs(bg[5] - bo[5]).max() / np.abs(bo[5]).max()))
        worst["Y"] = max(worst.get("Y", 0), float(np.abs(yg - yo).max()))
    print(f"[batched] mixed lattice after 10 steps: {worst}", flush=True)
    for k, v in worst.items():
        assert v <= 1e-4, (k, v)


def test_surrogate_execution_modes_identical():
    """The surrogate properties kernel walks the chunks of all blocks of a
    batch with a persistent grid and loads each workgroup's next chunk (often
    in another block) while the network runs on the current one: per-block,
    batched and graph-replayed steps on the lattice give identical bytes.
    The sampled Wilke check keeps the network's values (tolerance 1e30), as
    the wavefronts it samples depend on the launch."""
    from combustioncodegpu_amd.lib import Solver
    t = Solver(cases.c1(24, 20, 16))
    t.mlp_train(epochs=2)
    params = t.mlp_get_params()
    t.close()
    case = cases.lattice(SMALL, n_spec=9)
    case.params["precision"] = abi.MIXED
    nb = len(case.blocks)

    def setup(s):
        s.mlp_set_params(*params)
        s.mlp_set_check(1e30, 1)
        s.set_transport(abi.TRANSPORT_MLP)

    got = {}
    for name, batched, graph in (("per-block", False, False), ("batched", True, False), ("graph", True, True)):
        print(f"[batched] surrogate modes: {name}", flush=True)
        g = run(case, 3, batched=batched, graph=graph, setup=setup)
        st = state(g, nb)
        got[name] = [np.concatenate([st[b], g.download(b, abi.PROPS).ravel()]) for b in range(nb)]
        stats = g.mlp_stats()
        rc, nfail = g.check()
        assert rc == 0 and nfail == 0
        assert stats["checked"] > 0 and stats["check_failed"] == 0
        g.close()
    for name in ("batched", "graph"):
        for b in range(nb):
            assert np.array_equal(got[name][b], got["per-block"][b]), f"{name}: block {b} differs"
