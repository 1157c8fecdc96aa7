"""Synthetic QUBO instances of the benchmark configurations.

All generators are deterministic functions of (seed, instance id):
uniform and Gaussian coefficients come from Philox4x32-10 (key = (seed,
instance), counter = (element index, tag)), graph structure from numpy's
PCG64 seeded with (seed, instance).

* dense_uniform(n)       -- upper triangle incl. diagonal i.i.d. U[-1,1), fp32
* sk_qubo(n)             -- SK Ising J_ij ~ N(0,1)/sqrt(n), h=0, mapped to QUBO
* regular_csr(n, d)      -- random d-regular graph, weights +-1, diag U[-0.5,0.5]
* batch_c3(...)          -- 512 such instances with n log-uniform in [64, 2048]
"""
from __future__ import annotations

import numpy as np

_M0 = np.uint64(0xD2511F53)
_M1 = np.uint64(0xCD9E8D57)
_W0 = np.uint32(0x9E3779B9)
_W1 = np.uint32(0xBB67AE85)
_MASK = np.uint64(0xFFFFFFFF)

TAG_GEN_UNIFORM = 0xC1
TAG_GEN_GAUSS = 0xC2


def philox4x32_10(c0, c1, c2, c3, k0, k1):
    """Vectorised Philox4x32-10 (SC'11); arguments broadcast, returns 4 uint32 arrays."""
    c0, c1, c2, c3 = (np.asarray(c, dtype=np.uint32) for c in np.broadcast_arrays(c0, c1, c2, c3))
    k0 = np.uint32(k0)
    k1 = np.uint32(k1)
    for r in range(10):
        if r:
            k0 = np.uint32((int(k0) + int(_W0)) & 0xFFFFFFFF)
            k1 = np.uint32((int(k1) + int(_W1)) & 0xFFFFFFFF)
        p0 = _M0 * c0.astype(np.uint64)
        p1 = _M1 * c2.astype(np.uint64)
        n0 = (p1 >> np.uint64(32)).astype(np.uint32) ^ c1 ^ k0
        n2 = (p0 >> np.uint64(32)).astype(np.uint32) ^ c3 ^ k1
        c1 = (p1 & _MASK).astype(np.uint32)
        c3 = (p0 & _MASK).astype(np.uint32)
        c0, c2 = n0, n2
    return c0, c1, c2, c3


def _upper_index(n):
    iu, ju = np.triu_indices(n)
    return iu, ju, (iu.astype(np.uint64) * np.uint64(n) + ju.astype(np.uint64))


def dense_uniform(n: int, seed: int = 0, instance: int = 0) -> np.ndarray:
    """Symmetric fp32 matrix; Q_ij, i <= j, = (w >> 8) * 2^-23 - 1 from counter
    (i*n + j, 0, 0, TAG_GEN_UNIFORM) of key (seed, instance)."""
    iu, ju, idx = _upper_index(n)
    w, _, _, _ = philox4x32_10((idx & _MASK).astype(np.uint32), (idx >> np.uint64(32)).astype(np.uint32),
                               0, TAG_GEN_UNIFORM, seed & 0xFFFFFFFF, instance)
    v = ((w >> np.uint32(8)).astype(np.float32) * np.float32(2.0 ** -23)) - np.float32(1.0)
    q = np.zeros((n, n), dtype=np.float32)
    q[iu, ju] = v
    q[ju, iu] = v
    return q


def sk_couplings(n: int, seed: int = 0, instance: int = 0) -> np.ndarray:
    """J (symmetric fp32, zero diagonal), J_ij ~ N(0,1)/sqrt(n) by Box-Muller on
    the two uniforms of counter (i*n + j, 0, 0, TAG_GEN_GAUSS)."""
    iu, ju = np.triu_indices(n, 1)
    idx = iu.astype(np.uint64) * np.uint64(n) + ju.astype(np.uint64)
    a, b, _, _ = philox4x32_10((idx & _MASK).astype(np.uint32), (idx >> np.uint64(32)).astype(np.uint32),
                               0, TAG_GEN_GAUSS, seed & 0xFFFFFFFF, instance)
    u1 = ((a >> np.uint32(8)).astype(np.float64) + 1.0) * 2.0 ** -24  # (0, 1]
    u2 = (b >> np.uint32(8)).astype(np.float64) * 2.0 ** -24
    z = np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)
    j = np.zeros((n, n), dtype=np.float32)
    v = (z / np.sqrt(n)).astype(np.float32)
    j[iu, ju] = v
    j[ju, iu] = v
    return j


def sk_qubo(n: int, seed: int = 0, instance: int = 0):
    """SK Ising H(s) = sum_{i<j} J_ij s_i s_j, s = 2x - 1, as a QUBO:
    Q_ij = 4 J_ij (i != j), Q_ii = -2 sum_{j != i} J_ij, H = E_qubo(x) + c,
    c = sum_{i<j} J_ij.  Returns (Q fp32 symmetric, c)."""
    j = sk_couplings(n, seed, instance)
    q = (4.0 * j).astype(np.float32)
    row = j.astype(np.float64).sum(axis=1)
    q[np.arange(n), np.arange(n)] = (-2.0 * row).astype(np.float32)
    c = float(np.triu(j.astype(np.float64), 1).sum())
    return q, c


def _random_regular_pairs(n, d, rng, max_rounds=200000):
    """Random d-regular simple graph: configuration model, then random edge
    switches until no self-loop or multi-edge is left."""
    stubs = np.repeat(np.arange(n, dtype=np.int64), d)
    rng.shuffle(stubs)
    pairs = stubs.reshape(-1, 2).copy()
    m = len(pairs)
    for _ in range(max_rounds):
        u, v = pairs[:, 0], pairs[:, 1]
        key = np.minimum(u, v) * n + np.maximum(u, v)
        order = np.argsort(key, kind="stable")
        dup = np.zeros(m, dtype=bool)
        ks = key[order]
        dup[order[1:]] = ks[1:] == ks[:-1]
        bad = np.flatnonzero((u == v) | dup)
        if bad.size == 0:
            return pairs
        i = bad[0]
        k = int(rng.integers(m))
        if k == i:
            continue
        a, b = pairs[i]
        c, e = pairs[k]
        if rng.integers(2):
            c, e = e, c
        # (a,b),(c,e) -> (a,c),(b,e); accepted if it creates no loop
        if a != c and b != e:
            pairs[i] = (a, c)
            pairs[k] = (b, e)
    raise RuntimeError("random regular graph: repair did not converge")


def regular_csr(n: int, d: int, seed: int = 0, instance: int = 0):
    """Random d-regular graph on n nodes, weights uniform {-1,+1}, diagonal
    U[-0.5,0.5] (fp32); returns CSR (row_ptr int32, col int32, val fp32) of the
    symmetric off-diagonal couplings and diag fp32."""
    if (n * d) % 2:
        raise ValueError("n*d must be even")
    rng = np.random.Generator(np.random.PCG64([seed, instance]))
    pairs = _random_regular_pairs(n, d, rng)
    w = rng.choice(np.array([-1.0, 1.0], dtype=np.float32), size=len(pairs))
    diag = (rng.random(n, dtype=np.float32) - np.float32(0.5)).astype(np.float32)
    rows = np.concatenate([pairs[:, 0], pairs[:, 1]])
    cols = np.concatenate([pairs[:, 1], pairs[:, 0]])
    vals = np.concatenate([w, w])
    order = np.lexsort((cols, rows))
    rows, cols, vals = rows[order], cols[order], vals[order]
    row_ptr = np.zeros(n + 1, dtype=np.int32)
    np.add.at(row_ptr, rows + 1, 1)
    row_ptr = np.cumsum(row_ptr).astype(np.int32)
    return row_ptr, cols.astype(np.int32), vals.astype(np.float32), diag


def csr_to_dense(row_ptr, col, val, diag):
    n = len(diag)
    q = np.zeros((n, n), dtype=np.float64)
    rows = np.repeat(np.arange(n), np.diff(row_ptr))
    q[rows, col] = val
    q[np.arange(n), np.arange(n)] = diag
    return q


def batch_c3(num_instances: int = 512, seed: int = 0, n_min: int = 64, n_max: int = 2048):
    """Config-3 batch: n = round(exp U[ln n_min, ln n_max]) (bumped by one when
    n*d is odd), d uniform in {3..8}.  Returns a list of CSR tuples."""
    rng = np.random.Generator(np.random.PCG64([seed, 0xC3]))
    out = []
    for m in range(num_instances):
        n = int(round(np.exp(rng.uniform(np.log(n_min), np.log(n_max)))))
        d = int(rng.integers(3, 9))
        if (n * d) % 2:
            n += 1
        out.append(regular_csr(n, d, seed, m + 1))
    return out


def write_qubo(path: str, q: np.ndarray):
    """Write a symmetric matrix as a `.qubo` file (upper triangle incl. the
    diagonal, values with 9 significant digits)."""
    n = q.shape[0]
    iu, ju = np.triu_indices(n)
    nlin = n
    nquad = n * (n - 1) // 2
    with open(path, "w") as f:
        f.write(f"p qubo 0 {n} {nlin} {nquad}\n")
        for i, j in zip(iu, ju):
            f.write(f"{i} {j} {float(q[i, j]):.9g}\n")
