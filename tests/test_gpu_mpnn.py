"""GPU checks of the warm-start network's forward/backward kernels.

The oracle (oracle/mpnn.py, FP64, its analytic gradient pinned by finite
differences on the CPU) is run in bf16 mode, which rounds the operands the
GPU rounds; the remaining difference is fp32 accumulation and the rounding of
hidden states computed in fp32 instead of FP64, hence the tolerances below.
"""
import numpy as np
import pytest

import onesolver_amd
from oracle import mpnn

pytestmark = pytest.mark.gpu


def _data(ctx, I=3, n=64, seed=9):
    d = onesolver_amd.DenseBatch(ctx, sk=(I, n, seed, 0))
    return d, d.q()


def _params(seed):
    flat = mpnn.init_params(seed)
    P = mpnn.unpack(flat)
    for l in range(3):
        P[f"b{l}"] += 0.02 * np.arange(64) / 64.0
    P["bo"][0] = 0.1
    return mpnn.pack(P)


def test_mpnn_gradient_matches_oracle(ctx):
    d, qs = _data(ctx)
    rng = np.random.default_rng(1)
    labels = rng.integers(0, 2, (d.I, d.n), dtype=np.uint8)
    net = onesolver_amd.WarmStartNet(ctx, seed=3)
    assert net.num_params == mpnn.NUM_PARAMS
    flat = _params(4)
    net.set_params(flat)
    # parameters travel as fp32
    assert np.array_equal(net.params(), flat.astype(np.float32))
    print("[gpu] mpnn loss/grad vs oracle, 3 x n=64", flush=True)
    loss, g = net.loss_grad(d, labels)
    flat32 = flat.astype(np.float32).astype(np.float64)
    ol, og = mpnn.loss_grad(flat32, list(qs), list(mpnn.canonical(labels)), rounding="bf16")
    assert abs(loss - ol) < 1e-4 * abs(ol)
    at = 0
    for name, shape in mpnn.SHAPES:
        k = int(np.prod(shape))
        a, b = g[at:at + k].astype(np.float64), og[at:at + k]
        at += k
        err = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-12)
        assert err < 2e-2, (name, err)
    rel = np.linalg.norm(g - og) / np.linalg.norm(og)
    assert rel < 1e-2, rel
    net.close()
    d.close()


def test_mpnn_predict_matches_oracle(ctx):
    d, qs = _data(ctx, I=2, n=100, seed=2)
    net = onesolver_amd.WarmStartNet(ctx, seed=1)
    flat = _params(7)
    net.set_params(flat)
    print("[gpu] mpnn predict n=100 (padded rows)", flush=True)
    p = net.predict(d)
    assert p.shape == (2, 100)
    for i in range(2):
        op = mpnn.forward(flat.astype(np.float32).astype(np.float64), qs[i], rounding="bf16")
        assert np.max(np.abs(p[i] - op)) < 5e-3
    net.close()
    d.close()


def test_mpnn_training_step_is_deterministic_and_learns(ctx):
    d, qs = _data(ctx, I=4, n=128, seed=5)
    # labels with structure the network can express: spin i = 1 iff its
    # diagonal is negative (a function of the first feature)
    labels = (np.stack([np.diag(q) for q in qs]) < 0).astype(np.uint8)
    labels[:, 0] = 1
    a = onesolver_amd.WarmStartNet(ctx, seed=11)
    b = onesolver_amd.WarmStartNet(ctx, seed=11)
    assert np.array_equal(a.params(), b.params())
    print("[gpu] mpnn 60 Adam steps", flush=True)
    la = [a.train_step(d, labels, 3e-3) for _ in range(60)]
    lb = [b.train_step(d, labels, 3e-3) for _ in range(60)]
    assert la == lb  # deterministic reductions
    assert np.array_equal(a.params(), b.params())
    assert la[-1] < 0.5 * la[0], (la[0], la[-1])
    a.close()
    b.close()
    d.close()


def test_mpnn_rejects_bad_ranges(ctx):
    d, _ = _data(ctx, I=2, n=64)
    net = onesolver_amd.WarmStartNet(ctx)
    with pytest.raises(onesolver_amd.OneSolverError):
        net.predict(d, first=1, count=2)
    with pytest.raises(ValueError):
        net.set_params(np.zeros(5))
    net.close()
    d.close()
