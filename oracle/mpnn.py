"""CPU FP64 restatement of the warm-start message-passing network (test
infrastructure: only tests, smoke() and bench.py's CPU baseline use it).

The network has no counterpart in the reference; it is pinned here by
itself: the analytic gradient of ``loss_grad`` in exact mode is checked
against central finite differences of ``loss`` (tests/test_mpnn_oracle.py),
and the GPU kernels are checked against this module in "bf16" mode, which
rounds exactly the operands the GPU rounds to bf16 (the coupling matrix, the
weight matrices and the inputs of the two neighbour aggregations), so that
the remaining difference is fp32 accumulation.

Model (per instance, Q symmetric n x n):
  scale = sqrt(mean_i sum_{j!=i} Q_ij^2),  A = offdiag(Q) / scale,
  X0[i] = (Q_ii / scale, sum_{j!=i} Q_ij / scale, [i == 0], 1),
  for l = 1..3:  M = A r(H),  Z = H r(Ws_l) + M r(Wm_l) + b_l,  H = relu(Z)
  logit = H r(wo) + bo,  p = sigmoid(logit)
  loss = mean over nodes (and instances) of BCE(logit, y),
with r = bf16 rounding (bf16 mode) or the identity (exact mode).  Labels are
canonical: a state and its complement have equal SK energy, so states are
flipped to x_0 = 1 and spin 0 carries the pin feature.
"""
from __future__ import annotations

import numpy as np

F0 = 4
HID = 64
LAYERS = 3
SHAPES = []  # (name, shape) in the packed parameter order of the C ABI
for _l in range(LAYERS):
    _fin = F0 if _l == 0 else HID
    SHAPES += [(f"ws{_l}", (_fin, HID)), (f"wm{_l}", (_fin, HID)), (f"b{_l}", (HID,))]
SHAPES += [("wo", (HID,)), ("bo", (1,))]
NUM_PARAMS = int(sum(np.prod(s) for _, s in SHAPES))


def unpack(flat):
    out, at = {}, 0
    for name, shape in SHAPES:
        k = int(np.prod(shape))
        out[name] = np.asarray(flat[at:at + k], dtype=np.float64).reshape(shape)
        at += k
    return out


def pack(params):
    return np.concatenate([np.asarray(params[name], dtype=np.float64).ravel()
                           for name, _ in SHAPES])


def bf16(x):
    """Round to bf16 (round to nearest even, via fp32), returned as float64."""
    f = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + np.uint64(0x7FFF) + ((u >> np.uint64(16)) & np.uint64(1))) & np.uint64(0xFFFF0000)
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def features(q):
    """(A, X0) of one instance; A as the GPU stores it is bf16(fp32(A))."""
    q = np.asarray(q, dtype=np.float64)
    n = q.shape[0]
    off = q.copy()
    np.fill_diagonal(off, 0.0)
    scale = np.sqrt((off * off).sum(axis=1).mean())
    if scale == 0.0:
        scale = 1.0
    a = off / scale
    x0 = np.zeros((n, F0))
    x0[:, 0] = np.diag(q) / scale
    x0[:, 1] = off.sum(axis=1) / scale
    x0[0, 2] = 1.0
    x0[:, 3] = 1.0
    return a, x0


def canonical(states):
    """Flip every state whose spin 0 is 0 (the complement has the same SK energy)."""
    s = np.array(states, dtype=np.uint8, copy=True)
    flip = s[..., 0] == 0
    s[flip] = 1 - s[flip]
    return s


def _forward(P, a, x0, r):
    cache = []
    h = x0
    for l in range(LAYERS):
        m = a @ r(h)
        z = h @ r(P[f"ws{l}"]) + m @ r(P[f"wm{l}"]) + P[f"b{l}"]
        cache.append((h, m, z))
        h = np.maximum(z, 0.0)
    logit = h @ r(P["wo"]) + P["bo"][0]
    return logit, h, cache


def forward(flat, q, rounding="exact"):
    """Marginals p (n,) of one instance."""
    r = bf16 if rounding == "bf16" else (lambda v: v)
    P = unpack(flat)
    a, x0 = features(q)
    if rounding == "bf16":
        a = bf16(a)
    logit, _, _ = _forward(P, a, x0, r)
    return 1.0 / (1.0 + np.exp(-logit))


def _bce(logit, y):
    return np.maximum(logit, 0.0) - logit * y + np.log1p(np.exp(-np.abs(logit)))


def loss(flat, qs, labels, rounding="exact"):
    r = bf16 if rounding == "bf16" else (lambda v: v)
    P = unpack(flat)
    total, count = 0.0, 0
    for q, y in zip(qs, labels):
        a, x0 = features(q)
        if rounding == "bf16":
            a = bf16(a)
        logit, _, _ = _forward(P, a, x0, r)
        total += _bce(logit, np.asarray(y, dtype=np.float64)).sum()
        count += len(y)
    return total / count


def loss_grad(flat, qs, labels, rounding="exact"):
    """Mean BCE over all nodes of all instances and its gradient (packed)."""
    r = bf16 if rounding == "bf16" else (lambda v: v)
    P = unpack(flat)
    G = {name: np.zeros(shape) for name, shape in SHAPES}
    count = sum(len(y) for y in labels)
    total = 0.0
    for q, y in zip(qs, labels):
        y = np.asarray(y, dtype=np.float64)
        a, x0 = features(q)
        if rounding == "bf16":
            a = bf16(a)
        logit, h3, cache = _forward(P, a, x0, r)
        total += _bce(logit, y).sum()
        p = 1.0 / (1.0 + np.exp(-logit))
        dlogit = (p - y) / count
        G["wo"] += h3.T @ dlogit
        G["bo"][0] += dlogit.sum()
        dh = np.outer(dlogit, r(P["wo"]))
        for l in reversed(range(LAYERS)):
            h, m, z = cache[l]
            dz = dh * (z > 0.0)
            G[f"b{l}"] += dz.sum(axis=0)
            G[f"ws{l}"] += h.T @ dz
            G[f"wm{l}"] += m.T @ dz
            if l > 0:
                dh = dz @ r(P[f"ws{l}"]).T + a @ r(dz @ r(P[f"wm{l}"]).T)
    return total / count, pack(G)


def init_params(seed=0):
    """Glorot-uniform weights, zero biases (the library's own initialiser is
    checked for its shape only; values come from its Philox stream)."""
    rng = np.random.default_rng(seed)
    P = {}
    for name, shape in SHAPES:
        if len(shape) == 2:
            lim = np.sqrt(6.0 / (shape[0] + shape[1]))
            P[name] = rng.uniform(-lim, lim, shape)
        elif name == "wo":
            lim = np.sqrt(6.0 / (HID + 1))
            P[name] = rng.uniform(-lim, lim, shape)
        else:
            P[name] = np.zeros(shape)
    return pack(P)
