"""The FP64 network oracle pinned by finite differences and closed forms (CPU)."""
import numpy as np

from oracle import mpnn


def _instances(k, n, seed):
    rng = np.random.default_rng(seed)
    qs, ys = [], []
    for _ in range(k):
        q = rng.normal(size=(n, n))
        q = np.triu(q) + np.triu(q, 1).T
        qs.append(q)
        ys.append(rng.integers(0, 2, n))
    return qs, ys


def test_parameter_layout():
    assert mpnn.NUM_PARAMS == 2 * 4 * 64 + 64 + 2 * (2 * 64 * 64 + 64) + 64 + 1
    flat = mpnn.init_params(3)
    assert flat.shape == (mpnn.NUM_PARAMS,)
    assert np.array_equal(mpnn.pack(mpnn.unpack(flat)), flat)


def test_bf16_rounding_is_round_to_nearest_even():
    x = np.array([1.0, 1 + 2 ** -8, 1 + 3 * 2 ** -8, 1 + 2 ** -7, -2.5, 3.0e38, 1e-40])
    got = mpnn.bf16(x)
    assert got[0] == 1.0
    assert got[1] == 1.0                      # tie -> even
    assert got[2] == 1 + 2 ** -6              # tie -> even
    assert got[3] == 1 + 2 ** -7              # representable
    assert got[4] == -2.5
    assert abs(got[5] - 3.0e38) / 3.0e38 < 2 ** -8


def test_features_closed_form():
    q = np.array([[2.0, 3.0, 0.0], [3.0, -1.0, 4.0], [0.0, 4.0, 0.5]])
    a, x0 = mpnn.features(q)
    scale = np.sqrt((9 + 25 + 16) / 3.0)
    assert np.allclose(a * scale, [[0, 3, 0], [3, 0, 4], [0, 4, 0]])
    assert np.allclose(x0[:, 0] * scale, [2.0, -1.0, 0.5])
    assert np.allclose(x0[:, 1] * scale, [3.0, 7.0, 4.0])
    assert list(x0[:, 2]) == [1, 0, 0] and list(x0[:, 3]) == [1, 1, 1]


def test_canonical_labels_fix_spin_zero():
    s = np.array([[0, 1, 1], [1, 0, 0]], dtype=np.uint8)
    c = mpnn.canonical(s)
    assert c.tolist() == [[1, 0, 0], [1, 0, 0]]


def test_analytic_gradient_matches_finite_differences():
    qs, ys = _instances(2, 10, 0)
    flat = mpnn.init_params(1)
    P = mpnn.unpack(flat)
    for l in range(3):
        P[f"b{l}"] += 0.05  # keep pre-activations away from the relu kink
    flat = mpnn.pack(P)
    loss, grad = mpnn.loss_grad(flat, qs, ys)
    assert abs(loss - mpnn.loss(flat, qs, ys)) < 1e-14
    rng = np.random.default_rng(5)
    idx = np.concatenate([rng.choice(mpnn.NUM_PARAMS, 60, replace=False),
                          [mpnn.NUM_PARAMS - 1, mpnn.NUM_PARAMS - 2, 0, 300]])
    eps = 1e-6
    for i in idx:
        e = np.zeros_like(flat)
        e[i] = eps
        fd = (mpnn.loss(flat + e, qs, ys) - mpnn.loss(flat - e, qs, ys)) / (2 * eps)
        assert abs(fd - grad[i]) <= 1e-7 + 1e-5 * abs(fd), (i, fd, grad[i])


def _mutant_grad(flat, qs, ys):
    """mpnn.loss_grad with a plausible mistake planted: the neighbour term of
    the hidden-state gradient (A @ (dZ Wm^T)) dropped."""
    P = mpnn.unpack(flat)
    G = {name: np.zeros(shape) for name, shape in mpnn.SHAPES}
    count = sum(len(y) for y in ys)
    for q, y in zip(qs, ys):
        a, x0 = mpnn.features(q)
        logit, h3, cache = mpnn._forward(P, a, x0, lambda v: v)
        dlogit = (1.0 / (1.0 + np.exp(-logit)) - y) / count
        G["wo"] += h3.T @ dlogit
        G["bo"][0] += dlogit.sum()
        dh = np.outer(dlogit, P["wo"])
        for l in reversed(range(3)):
            h, m, z = cache[l]
            dz = dh * (z > 0.0)
            G[f"b{l}"] += dz.sum(axis=0)
            G[f"ws{l}"] += h.T @ dz
            G[f"wm{l}"] += m.T @ dz
            if l > 0:
                dh = dz @ P[f"ws{l}"].T  # mistake: "+ a @ (dz @ P[wm].T)" missing
    return mpnn.pack(G)


def test_planted_backward_mistake_is_caught():
    qs, ys = _instances(2, 10, 2)
    flat = mpnn.init_params(4)
    bad = _mutant_grad(flat, qs, ys)
    _, good = mpnn.loss_grad(flat, qs, ys)
    eps = 1e-6
    fails = 0
    for i in range(0, 2 * 4 * 64, 7):  # layer-1 weights see the dropped term
        e = np.zeros_like(flat)
        e[i] = eps
        fd = (mpnn.loss(flat + e, qs, ys) - mpnn.loss(flat - e, qs, ys)) / (2 * eps)
        assert abs(fd - good[i]) <= 1e-7 + 1e-5 * abs(fd)
        fails += abs(fd - bad[i]) > 1e-7 + 1e-5 * abs(fd)
    assert fails > 10


def test_bf16_mode_is_close_to_exact():
    qs, ys = _instances(2, 16, 3)
    flat = mpnn.init_params(2)
    l1, g1 = mpnn.loss_grad(flat, qs, ys)
    l2, g2 = mpnn.loss_grad(flat, qs, ys, rounding="bf16")
    assert abs(l1 - l2) < 1e-2
    assert np.linalg.norm(g1 - g2) < 0.15 * np.linalg.norm(g1)
