"""Cases read from reference-format input decks, stepped on the GPU and by
the oracle: the deck path hands the library the same descriptors and initial
field as any other caller, and the step keeps the FP64 parity bar (1e-12 per
step, normwise per quantity)."""
import pytest

from combustioncodegpu_amd import deck
from tests import deck_cases
from tests.test_gpu_parity import compare_state

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("which,steps", [("walls", 3), ("lattice", 2)])
def test_deck_case_parity(tmp_path, which, steps):
    from combustioncodegpu_amd.lib import Solver
    from oracle.pyoracle import Oracle
    make = deck_cases.walls_deck if which == "walls" else deck_cases.lattice_deck
    make(str(tmp_path))
    case = deck.load_case(str(tmp_path))
    print(f"[deck] {which} blocks={[(b.mx, b.my, b.mz) for b in case.blocks]} steps={steps}", flush=True)
    o, g = Oracle(case), Solver(case)
    for c in (o, g):
        c.upload_initial()
        c.init_conservative()
    compare_state(case, o, g, 1e-12, "init")
    o.step(steps)
    g.step(steps)
    compare_state(case, o, g, 1e-12 * steps, f"{steps} steps")
    rc, nfail = g.check()
    assert rc == 0 and nfail == 0


def test_context_coordinates_match_host_ghost_grid():
    """CFX_COORD of a context (the ghost grid its metrics were computed from)
    equals the host-only cfx_ghost_grid, block by block."""
    import numpy as np
    from combustioncodegpu_amd import abi, cases, lib
    case = cases.lattice([[7, 6], [6, 5], [5, 6]])
    print("[deck] coordinates of a 2x2x2 lattice", flush=True)
    host = lib.ghost_grid(case)
    g = lib.Solver(case)
    for b in range(len(case.blocks)):
        assert np.array_equal(g.download(b, abi.COORD), host[b])
    g.close()


def test_deck_case_mixed_graph(tmp_path):
    """The deck-loaded walls case (two initial groups) in mixed precision,
    batched and graph-captured, against the FP64 oracle after 5 steps: the
    mixed-precision error bar (normwise 1e-5 per quantity here)."""
    from combustioncodegpu_amd import abi
    from combustioncodegpu_amd.lib import Solver
    from oracle.pyoracle import Oracle
    from tests.helpers import assert_close_sel, interior
    deck_cases.walls_deck(str(tmp_path))
    ref = deck.load_case(str(tmp_path))
    mixed = deck.load_case(str(tmp_path), precision=abi.MIXED)
    print("[deck] walls case, mixed + graph, 5 steps", flush=True)
    o, g = Oracle(ref), Solver(mixed)
    g.set_execution(batched=True, graph=True)
    for c in (o, g):
        c.upload_initial()
        c.init_conservative()
        c.step(5)
    for sel in (abi.CV, abi.BV):
        assert_close_sel(sel, interior(g.download(0, sel)), interior(o.download(0, sel)), 1e-5, f"mixed {sel}")
    rc, nfail = g.check()
    assert rc == 0 and nfail == 0
    g.close()
