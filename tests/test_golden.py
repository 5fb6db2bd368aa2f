"""Committed golden fixtures (tests/golden/*.npz, written by
tests/golden/make_golden.py with the oracle): the seeded generators still
produce the same inputs, the oracle still produces the same steps, and the
HIP path matches them at the FP64 step bar (1e-12 per step)."""
import glob
import os

import numpy as np
import pytest

from combustioncodegpu_amd import abi
from tests.golden.make_golden import FIXTURES, input_digest, run_oracle
from tests.helpers import assert_close_sel

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
NAMES = sorted(FIXTURES)


def load(name):
    return dict(np.load(os.path.join(HERE, name + ".npz"), allow_pickle=False))


def test_every_fixture_is_listed():
    on_disk = sorted(os.path.splitext(os.path.basename(p))[0] for p in glob.glob(os.path.join(HERE, "*.npz")))
    assert on_disk == NAMES


@pytest.mark.parametrize("name", NAMES)
def test_generator_inputs_unchanged(name):
    make, _ = FIXTURES[name]
    assert str(load(name)["digest"]) == input_digest(make())


@pytest.mark.parametrize("name", NAMES)
def test_oracle_reproduces_fixture(name):
    make, steps = FIXTURES[name]
    gold = load(name)
    assert int(gold["steps"]) == steps
    out = run_oracle(make(), steps)
    for k, v in out.items():
        sel = abi.CV if k.startswith("cv") else abi.BV
        assert_close_sel(sel, v, gold[k], 1e-13, f"{name} {k}")


@pytest.mark.gpu
@pytest.mark.parametrize("name", NAMES)
def test_gpu_matches_fixture(name):
    from combustioncodegpu_amd.lib import Solver
    make, steps = FIXTURES[name]
    gold = load(name)
    case = make()
    print(f"[golden] {name}", flush=True)
    g = Solver(case)
    g.upload_initial()
    g.init_conservative()
    g.step(steps)
    for b in range(len(case.blocks)):
        for key, sel in ((f"cv{b}", abi.CV), (f"bv{b}", abi.BV)):
            got = g.download(b, sel)[..., 2:-2, 2:-2, 2:-2]
            assert_close_sel(sel, got, gold[key], 1e-12 * steps, f"{name} {key}")
