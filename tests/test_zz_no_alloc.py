"""cfx_mlp_predict_noalloc (the surrogate inference entry point that
allocates nothing) against cfx_mlp_predict through the C ABI: two calls in a
row, the second with fewer points, must equal the allocating entry point's
outputs exactly (same kernel, same operands); a call over the array's limit
is rejected."""
import numpy as np
import pytest

from combustioncodegpu_amd import abi, cases
from tests.test_oracle_mlp import random_params, samples

pytestmark = pytest.mark.gpu

NS = 9
NIN = NOUT = NS + 2


def test_noalloc_predict_equals_predict():
    from combustioncodegpu_amd.driver import CfxError
    from combustioncodegpu_amd.lib import Solver
    c = cases.c1(16, 12, 10)
    c.params["precision"] = abi.MIXED
    g = Solver(c)
    norm = np.zeros(64, dtype=np.float32)
    norm[16:32] = 1.0
    norm[48:64] = 1.0
    g.mlp_set_params(random_params(NIN, NOUT, 21), norm)
    for n, seed in ((5000, 31), (1203, 32)):
        x, _ = samples(n, NIN, NOUT, seed)
        print(f"[noalloc] n={n}", flush=True)
        ref = g.mlp_predict(x)
        got = g.mlp_predict_noalloc(x)
        assert got.shape == ref.shape
        assert np.array_equal(got, ref), np.abs(got - ref).max()
        assert np.all(got[:, NOUT:] == 0.0)
    x, _ = samples(65537, NIN, NOUT, 33)
    with pytest.raises(CfxError):
        g.mlp_predict_noalloc(x)
