"""The transport surrogate on the GPU.

  - forward pass (BF16 MFMA, FP32 accumulation, the operand lane layout)
    against the oracle's network with the same BF16 rounding points;
  - gradients of the GPU training kernels against the oracle's
    backpropagation (same rounding points), and against exact FP64;
  - training on samples of the FP64 Wilke routine reaches a stated accuracy
    on held-out samples;
  - what the surrogate gains over the method without it: the same
    mixed-precision step with Wilke transport, timed per kernel class, and
    the surrogate step's error against the FP64 oracle after 10 steps.
"""
import numpy as np
import pytest

from combustioncodegpu_amd import abi, cases
from tests.test_oracle_mlp import random_params, samples

pytestmark = pytest.mark.gpu

NS = 9
NIN = NOUT = NS + 2


def identity_norm():
    n = np.zeros(64, dtype=np.float32)
    n[16:32] = 1.0  # input 1/std
    n[48:64] = 1.0  # output std
    return n


def solver(case):
    from combustioncodegpu_amd.lib import Solver
    return Solver(case)


def small_mixed_case(mlp=False):
    c = cases.c1(24, 20, 16)
    c.params["precision"] = abi.MIXED
    if mlp:
        c.params["transport"] = abi.TRANSPORT_MLP
    return c


def test_forward_matches_oracle():
    from oracle import pyoracle
    g = solver(small_mixed_case())
    p = random_params(NIN, NOUT, 11)
    g.mlp_set_params(p, identity_norm())
    x, _ = samples(1000, NIN, NOUT, 12)
    y = g.mlp_predict(x)[:, :NOUT]
    ye = pyoracle.mlp_forward(p, NIN, NOUT, x, emulate=True)[:, :NOUT]
    yx = pyoracle.mlp_forward(p, NIN, NOUT, x, emulate=False)[:, :NOUT]
    err_e = np.abs(y - ye).max() / np.abs(ye).max()
    err_x = np.abs(y - yx).max() / np.abs(yx).max()
    print(f"[mlp] forward: vs emulated oracle {err_e:.2e}, vs FP64 {err_x:.2e}", flush=True)
    assert err_e <= 2e-3
    assert err_x <= 5e-2


def test_gradient_matches_oracle():
    from oracle import pyoracle
    g = solver(small_mixed_case())
    p = random_params(NIN, NOUT, 21)
    g.mlp_set_params(p, identity_norm())
    x, t = samples(256, NIN, NOUT, 22)
    grad, loss = g.mlp_gradient(x, t)
    le, ge = pyoracle.mlp_loss_grad(p, NIN, NOUT, x, t, emulate=True)
    lx, gx = pyoracle.mlp_loss_grad(p, NIN, NOUT, x, t, emulate=False)
    err_e = np.abs(grad - ge).max() / np.abs(ge).max()
    err_x = np.abs(grad - gx).max() / np.abs(gx).max()
    print(f"[mlp] gradient: vs emulated oracle {err_e:.2e}, vs FP64 {err_x:.2e}; loss {loss:.6e} vs {le:.6e}",
          flush=True)
    assert abs(loss - le) <= 1e-3 * le
    assert err_e <= 5e-3
    assert err_x <= 5e-2


def test_training_reaches_accuracy():
    """The default training (2M samples of the FP64 Wilke routine) on the
    held-out samples the surrogate answers (inside the envelope, whose radius
    is the 0.99 quantile of the training inputs' Mahalanobis distance): at
    least 98% of them answered, 99% of mu, kappa and rho*D within 3 %, every
    one within 10 % (measured 1.8 %, 2.4 %, 5.3 %; with the envelope at the
    largest training distance the dominant species' rho*D reached 60 % in
    the sparse tail)."""
    g = solver(small_mixed_case())
    rep = g.mlp_train()
    print(f"[mlp] training: {rep}", flush=True)
    assert rep["n_train"] + rep["n_val"] == 2000000
    assert rep["loss_last"] < rep["loss_first"] / 1000
    assert rep["val_answered"] >= 0.98
    assert max(rep["val_p99_rel"]) <= 0.03
    assert max(rep["val_max_rel"]) <= 0.10
    rep_params, norm = g.mlp_get_params()
    assert np.all(np.isfinite(rep_params))


def test_surrogate_step_gain_and_error():
    """What the surrogate does to the mixed-precision step, against the same
    step with the Wilke routine: the error against the FP64 oracle after 10
    and 100 SSP-RK3 steps (bar 1e-4 per field), and the time of the
    properties pass.  On MI355X the surrogate's properties pass is slower
    than FP32 Wilke (FP32 Wilke with hardware rcp/sqrt/exp2/log2 already runs
    at the HBM floor of a separate properties pass, and the network adds 192
    SiLU evaluations per point); the test pins that documented loss on a
    block large enough that launch overhead does not hide it:
    1.05 <= t_props(surrogate) / t_props(Wilke) <= 2.0."""
    from oracle.pyoracle import Oracle
    ref = cases.c1(24, 20, 16)
    o = Oracle(ref)
    o.upload_initial()
    o.init_conservative()
    marks = (10, 100)
    ref_states = {}
    done = 0
    for m in marks:
        o.step(m - done)
        done = m
        ref_states[m] = (o.download(0, abi.BV), o.download(0, abi.YK))
    out = {}
    params = None
    for name in ("wilke", "mlp"):
        g = solver(small_mixed_case())
        if name == "mlp":
            g.mlp_train()
            params = g.mlp_get_params()
            g.set_transport(abi.TRANSPORT_MLP)
        g.upload_initial()
        g.init_conservative()
        errs, done = {}, 0
        for m in marks:
            g.step(m - done)
            done = m
            bo, yo = ref_states[m]
            bg, yg = g.download(0, abi.BV), g.download(0, abi.YK)
            vel = np.sqrt((bo[1:4] ** 2).sum(axis=0)).max()
            errs[m] = {"rho": float(np.abs(bg[0] - bo[0]).max() / np.abs(bo[0]).max()),
                       "T": float(np.abs(bg[5] - bo[5]).max() / np.abs(bo[5]).max()),
                       "Y": float(np.abs(yg - yo).max()),
                       "velocity": float(np.sqrt(((bg[1:4] - bo[1:4]) ** 2).sum(axis=0)).max() / vel)}
        out[name] = {"err": errs, "stats": g.mlp_stats() if name == "mlp" else None}
        rc, nfail = g.check()
        assert rc == 0 and nfail == 0
        g.close()
    # properties-pass time on a 128x64x64 block, trained network loaded
    t = {}
    for name in ("wilke", "mlp"):
        c = cases.c1(128, 64, 64)
        c.params["precision"] = abi.MIXED
        if name == "mlp":
            c.params["transport"] = abi.TRANSPORT_MLP
        g = solver(c)
        if name == "mlp":
            g.mlp_set_params(*params)
            g.set_transport(abi.TRANSPORT_MLP)
        g.upload_initial()
        g.init_conservative()
        g.step(2)
        g.set_profiling(True)
        g.time_steps(5)
        kt = g.kernel_times()
        t[name] = kt["props"][0] / 5
        g.close()
    ratio = t["mlp"] / t["wilke"]
    print(f"[mlp] ablation: {out}; props ms per step on 128x64x64: {t}, surrogate / Wilke {ratio:.3f}", flush=True)
    for m in marks:
        for k, v in out["mlp"]["err"][m].items():
            assert v <= 1e-4, (m, k, v)
    st = out["mlp"]["stats"]
    assert st["checked"] > 0 and st["check_failed"] <= st["checked"] // 100 + 1
    assert 1.05 <= ratio <= 2.0, ratio
