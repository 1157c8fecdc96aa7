"""`one-solver-anneal --mode sweep --device-type gpu`: the fast path behind the
drop-in CLI (sa::anneal_sweeps).

The CLI builds the reference's beta schedule and runs one instance as stream
0 of the dense or sparse path; its Solution CSV must hold the oracle's best
replica (lowest FP64 energy, first replica on ties) of the same chains, with
the energy printed to 6 significant digits.  With --warm-start the replicas
start from the network's marginals, as the library's own warm start does.
"""
import os
import subprocess

import numpy as np
import pytest

import onesolver_amd
import oracle
from onesolver_amd import instances
from qubo_corpus import dense_text

pytestmark = pytest.mark.gpu

CLI = onesolver_amd.CLI_PATH


def _run(args):
    r = subprocess.run([CLI] + args, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    return r.stdout


def _read_csv(path, n):
    with open(path) as f:
        header, row = f.read().splitlines()
    assert header == ",".join(str(i) for i in range(n)) + ",energy"
    t = row.split(",")
    return np.array([int(v) for v in t[:n]], dtype=np.uint8), float(t[n])


def _sparse_text(q):
    n = q.shape[0]
    iu, ju = np.nonzero(np.triu(q, 1))
    lines = ["p qubo 0 %d %d %d\n" % (n, n, len(iu))]
    lines += ["%d %d %.17g\n" % (i, i, q[i, i]) for i in range(n)]
    lines += ["%d %d %.17g\n" % (i, j, q[i, j]) for i, j in zip(iu, ju)]
    return "".join(lines)


def _best(be):
    return int(np.argmin(be))  # first index on ties


def _energy_printed(e, printed):
    tol = 0.5 * 10 ** (np.floor(np.log10(abs(e))) - 5)
    return abs(e - printed) <= tol * 1.0001


@pytest.mark.parametrize("beta_kind", ["inverse", "temperature"])
def test_cli_sweep_dense_path_is_the_oracle_best_replica(tmp_path, beta_kind):
    n, R, iters = 200, 64, 24
    rng = np.random.default_rng(3)
    q = np.triu(rng.integers(-3, 4, (n, n)).astype(np.float64))
    q += np.triu(q, 1).T
    src = tmp_path / "d.qubo"
    src.write_text(dense_text(q))
    out = tmp_path / "d.csv"
    bmin, bmax = (0.1, 3.0) if beta_kind == "inverse" else (0.3, 5.0)
    print(f"[gpu] CLI sweep dense n={n} {beta_kind}", flush=True)
    log = _run(["--input", str(src), "--output", str(out), "--mode", "sweep", "--device-type",
                "gpu", "--num-iter", str(iters), "--num-tries", str(R), "--beta-min", str(bmin),
                "--beta-max", str(bmax), "--beta-kind", beta_kind, "--seed", "5",
                "--sweeps-per-beta", "2"])
    assert "Mode: sweep (dense path" in log
    beta = np.repeat(oracle.geometric_schedule(bmin, bmax, iters), 2)
    ts = onesolver_amd.tscale_from_beta(beta, temperature=beta_kind == "temperature")
    _, bs, be = oracle.sweep_dense(q, n, ts, R, seed=5, instance=0)
    x, e = _read_csv(out, n)
    i = _best(be)
    assert (x == bs[i]).all()
    assert _energy_printed(be[i], e)


def test_cli_sweep_sparse_path_is_the_oracle_best_replica(tmp_path):
    n, d, R, iters = 300, 3, 96, 30
    rp, col, val, diag = instances.regular_csr(n, d, seed=4, instance=0)
    diag = (np.round(diag * 4) / 4).astype(np.float32)  # quarter steps: exact fp32 sums
    q = instances.csr_to_dense(rp, col, val, diag)
    src = tmp_path / "s.qubo"
    src.write_text(_sparse_text(q))
    out = tmp_path / "s.csv"
    print(f"[gpu] CLI sweep sparse n={n}", flush=True)
    log = _run(["--input", str(src), "--output", str(out), "--mode", "sweep", "--device-type",
                "gpu", "--num-iter", str(iters), "--num-tries", str(R), "--beta-max", "3",
                "--beta-kind", "inverse", "--seed", "9", "--schedule-type", "linear"])
    assert "Mode: sweep (sparse path" in log
    ts = onesolver_amd.tscale_from_beta(oracle.linear_schedule(0.1, 3.0, iters))
    fin, bs, be = oracle.sweep_csr(rp, col, val, diag, ts, R, seed=9, instance=0)
    e64 = oracle.energy(q, bs)
    x, e = _read_csv(out, n)
    i = _best(e64)
    assert (x == bs[i]).all()
    assert _energy_printed(e64[i], e)


def test_cli_sweep_warm_start_equals_the_library_warm_start(ctx, tmp_path):
    n, R, iters = 128, 64, 16
    data = onesolver_amd.DenseBatch(ctx, sk=(1, n, 21, 0))
    q = data.q()[0]
    net = onesolver_amd.WarmStartNet(ctx, seed=77)
    params = tmp_path / "net.bin"
    net.params().astype("<f4").tofile(params)
    p = net.predict(data)
    ts = onesolver_amd.tscale_from_beta(oracle.geometric_schedule(1.0, 3.0, iters))
    ref = data.anneal(R, ts, seed=1234, instance=0, init_marginals=p)
    net.close()
    data.close()
    src = tmp_path / "w.qubo"
    src.write_text(dense_text(q))
    out = tmp_path / "w.csv"
    print(f"[gpu] CLI sweep warm start n={n}", flush=True)
    _run(["--input", str(src), "--output", str(out), "--mode", "sweep", "--device-type", "gpu",
          "--num-iter", str(iters), "--num-tries", str(R), "--beta-min", "1", "--beta-max", "3",
          "--beta-kind", "inverse", "--warm-start", str(params)])
    x, e = _read_csv(out, n)
    assert (x == ref["best_state"][0]).all()
    assert _energy_printed(ref["best_energy"][0], e)
    # a parameter file of the wrong size is rejected with the reference's error exit
    bad = tmp_path / "bad.bin"
    bad.write_bytes(b"\0" * 12)
    r = subprocess.run([CLI, "--input", str(src), "--output", str(out), "--mode", "sweep",
                        "--device-type", "gpu", "--warm-start", str(bad)], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 1 and "fp32 parameters" in r.stderr
