"""GPU parity of one-solver-batch: each instance's CSV holds the best state
over its replicas and its FP64 energy, the same as the oracle's sweeps of
that instance alone with stream k (or_sweep_csr / or_sweep_dense), written
in the reference's Solution CSV format (include/model/solution.hpp:58-67)."""
import os
import subprocess

import numpy as np
import pytest

import onesolver_amd
import oracle
import qubo_corpus as qc

pytestmark = pytest.mark.gpu


def _run(args):
    r = subprocess.run([onesolver_amd.BATCH_CLI_PATH] + args, capture_output=True, text=True,
                       timeout=120)
    return r


def _read_csv(path):
    head, row = open(path).read().splitlines()
    vals = row.split(",")
    n = len(head.split(",")) - 1
    return np.array([int(v) for v in vals[:n]], np.uint8), vals[n]


def _csr(res):
    """Rows sorted by column, symmetric off-diagonal, and the diagonal."""
    n = res["n"]
    i, j, v = res["i"], res["j"], res["v"]
    off = i != j
    rows = np.concatenate([i[off], j[off]])
    cols = np.concatenate([j[off], i[off]])
    vals = np.concatenate([v[off], v[off]])
    order = np.lexsort((cols, rows))
    rows, cols, vals = rows[order], cols[order], vals[order]
    rp = np.zeros(n + 1, np.int32)
    np.add.at(rp, rows + 1, 1)
    diag = np.zeros(n)
    diag[i[~off]] = v[~off]
    return np.cumsum(rp).astype(np.int32), cols.astype(np.int32), vals.astype(np.float64), diag


def _sparse_text(n, d, seed):
    rng = np.random.default_rng(seed)
    edges = set()
    while len(edges) < n * d // 2:
        a, b = sorted(rng.integers(0, n, 2))
        if a != b:
            edges.add((int(a), int(b)))
    lines = ["%d %d %d\n" % (k, k, int(rng.integers(-2, 3))) for k in range(n)]
    lines += ["%d %d %d\n" % (a, b, int(rng.choice([-1, 1]))) for a, b in sorted(edges)]
    return "p qubo 0 %d %d %d\n" % (n, n, len(edges)) + "".join(lines)


@pytest.mark.parametrize("order", ["sequential", "colour"])
def test_batch_cli_sparse_instances_match_oracle(ctx, tmp_path, order):
    sizes = [(20, 3), (64, 4), (97, 5), (150, 3), (33, 6)]
    texts = [_sparse_text(n, d, 10 + k) for k, (n, d) in enumerate(sizes)]
    inp = tmp_path / "in"
    out = tmp_path / "out"
    inp.mkdir()
    out.mkdir()
    for k, t in enumerate(texts[:-1]):
        (inp / ("inst%02d.qubo" % k)).write_text(t)
    with onesolver_amd.Qubo(ctx, texts[-1]) as q:      # one instance as a binary file
        (inp / ("inst%02d.osqb" % (len(texts) - 1))).write_bytes(q.to_binary())
    S, R, seed = 40, 64, 7
    print(f"[gpu] one-solver-batch sparse, {order} order", flush=True)
    r = _run(["--input", str(inp), "--output", str(out), "--num-iter", str(S), "--num-tries",
              str(R), "--seed", str(seed), "--beta-max", "3", "--site-order", order])
    assert r.returncode == 0, r.stderr
    assert f"Path: sparse ({'colour-class' if order == 'colour' else 'sequential'} order)" in r.stdout
    ts = oracle.geometric_schedule(0.1, 3.0, S).astype(np.float32)   # beta as a temperature
    for k, t in enumerate(texts):
        _, _, res = oracle.qubo_parse(t.encode())
        rp, cl, vl, dg = _csr(res)
        od = oracle.csr_colour_order(rp, cl)[0] if order == "colour" else None
        _, bs, _ = oracle.sweep_csr(rp, cl, vl, dg, ts, R, seed=seed, instance=k, order=od)
        e64 = oracle.csr_energy(rp, cl, vl, dg, bs)
        b = int(np.argmin(e64))
        state, energy = _read_csv(out / ("inst%02d.csv" % k))
        assert (state == bs[b]).all(), k
        assert energy == "%g" % e64[b], (k, energy, e64[b])


def test_batch_cli_dense_manifest_matches_oracle(ctx, tmp_path):
    rng = np.random.default_rng(5)
    n, I = 48, 3
    qs = [np.triu(rng.integers(-3, 4, (n, n)).astype(np.float64)) for _ in range(I)]
    (tmp_path / "sub").mkdir()
    names = []
    for k, q in enumerate(qs):
        name = "sub/d%d.qubo" % k
        (tmp_path / name).write_text(qc.dense_text(q))
        names.append(name)
    (tmp_path / "list.txt").write_text("# three dense instances\n\n" + "\n".join(names) + "\n")
    out = tmp_path / "out"
    out.mkdir()
    S, R, seed = 30, 64, 3
    print("[gpu] one-solver-batch dense", flush=True)
    r = _run(["--input", str(tmp_path / "list.txt"), "--output", str(out), "--num-iter", str(S),
              "--num-tries", str(R), "--seed", str(seed), "--beta-kind", "inverse",
              "--beta-min", "0.2", "--beta-max", "2", "--schedule-type", "linear"])
    assert r.returncode == 0, r.stderr
    assert "Path: dense" in r.stdout
    beta = oracle.linear_schedule(0.2, 2.0, S)
    ts = (1.0 / beta).astype(np.float32)
    for k, q in enumerate(qs):
        flat = q + np.triu(q, 1).T
        _, bs, _ = oracle.sweep_dense(flat, n, ts, R, seed=seed, instance=k)
        e64 = oracle.energy(flat, bs)
        b = int(np.argmin(e64))
        state, energy = _read_csv(out / ("d%d.csv" % k))
        assert (state == bs[b]).all(), k
        assert energy == "%g" % e64[b], (k, energy, e64[b])


def test_batch_cli_reports_parse_errors(tmp_path):
    (tmp_path / "a.qubo").write_text("p qubo 0 4 1 0\n0 0 1\n")
    (tmp_path / "b.qubo").write_text("p qubo 0 100 2 2\n0 0 -0.5\n1 0 2.0\n")
    r = _run(["--input", str(tmp_path), "--output", str(tmp_path)])
    assert r.returncode == 1
    assert "b.qubo: Incorrect file, encountered coefficient from lower triangle" in r.stderr


def test_batch_cli_reports_the_first_failing_file_in_order(tmp_path):
    """The files are parsed by several host threads; with two bad files the
    run still stops with the message of the first in name order, as a serial
    run would, and writes no CSV."""
    good = _sparse_text(20, 3, 1)
    for k in range(12):
        text = good
        if k == 3:
            text = "p qubo 0 100 2 2\n0 0 -0.5\n1 0 2.0\n"
        if k == 8:
            text = "p qubo 0 4 1 0\nnot a coefficient line\n"
        (tmp_path / f"f{k:02d}.qubo").write_text(text)
    out = tmp_path / "out"
    out.mkdir()
    r = _run(["--input", str(tmp_path), "--output", str(out)])
    assert r.returncode == 1
    assert "f03.qubo: Incorrect file, encountered coefficient from lower triangle" in r.stderr
    assert "f08.qubo" not in r.stderr
    assert not list(out.iterdir())
    (tmp_path / "f03.qubo").write_text(good)
    r = _run(["--input", str(tmp_path), "--output", str(out)])
    assert r.returncode == 1 and "f08.qubo: " in r.stderr


def test_batch_cli_forced_dense_path_needs_one_size(tmp_path):
    (tmp_path / "a.qubo").write_text(_sparse_text(20, 3, 1))
    (tmp_path / "b.qubo").write_text(_sparse_text(30, 3, 2))
    r = _run(["--input", str(tmp_path), "--output", str(tmp_path), "--path", "dense",
              "--num-iter", "5", "--num-tries", "8"])
    assert r.returncode == 1
    assert "the dense path needs instances of one size" in r.stderr
