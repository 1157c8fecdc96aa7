"""GPU parity of the exhaustive search (os_exhaustive) with the oracle's
restatement of exhaustive::solve (include/exhaustive/exhaustive.hpp:29-167).

Both return the state of lowest literal FP64 energy (the reference's
summation order), lowest state index on ties, so states and energies are
compared exactly: on random instances against brute force, on the reference's
known-answer files, on the committed config-1 ground state, and on instances
built with exact ties.
"""
import json
import os
import subprocess

import numpy as np
import pytest

import onesolver_amd
import oracle
from conftest import GOLDEN
from qubo_io import load_qubo

pytestmark = pytest.mark.gpu


def _sym(q):
    """The symmetric matrix of q's upper triangle (the oracle's Gray-code
    search reads whole rows)."""
    return np.triu(q) + np.triu(q, 1).T


def _random_upper(n, seed, integer=False):
    rng = np.random.default_rng(seed)
    if integer:
        q = rng.integers(-2, 3, (n, n)).astype(np.float64)
    else:
        q = rng.uniform(-1, 1, (n, n))
    # entries below the diagonal are junk: the search reads the upper triangle
    return q


@pytest.mark.parametrize("n", [1, 2, 3, 7, 13, 20, 21, 23, 25])
def test_exhaustive_matches_oracle(ctx, n):
    q = _random_upper(n, 100 + n)
    print(f"[gpu] exhaustive n={n}", flush=True)
    st, e = ctx.exhaustive(q)
    ref_st, ref_e = oracle.exhaustive(_sym(q), n, literal=n <= 20)
    assert e == ref_e
    assert (st == ref_st).all()
    assert oracle.energy(_sym(q), st[None, :])[0] == e


@pytest.mark.parametrize("n", [18, 24, 26])
def test_exhaustive_integer_instances_with_ties(ctx, n):
    q = _random_upper(n, 7 * n, integer=True)
    print(f"[gpu] exhaustive integer n={n}", flush=True)
    st, e = ctx.exhaustive(q)
    ref_st, ref_e = oracle.exhaustive(_sym(q), n, literal=n <= 20)
    assert e == ref_e
    assert (st == ref_st).all()


def test_exhaustive_kats(ctx):
    with open(os.path.join(GOLDEN, "kat", "expected.json")) as f:
        kats = json.load(f)
    for k in kats:
        q = load_qubo(os.path.join(GOLDEN, "kat", k["file"]))
        n = q.shape[0]
        st, e = ctx.exhaustive(q)
        ref_st, ref_e = oracle.exhaustive(_sym(q), n, literal=True)
        assert e == ref_e and (st == ref_st).all(), k["file"]
        assert abs(e - k["energy"]) <= 1e-13  # tests/exhaustive_test.cpp:156


def test_exhaustive_config1_ground_state(ctx):
    q = load_qubo(os.path.join(GOLDEN, "c1_dense28.qubo"))
    with open(os.path.join(GOLDEN, "c1_ground.json")) as f:
        g = json.load(f)
    print("[gpu] exhaustive config 1, n=28", flush=True)
    st, e = ctx.exhaustive(q)
    assert repr(e) == g["energy_repr"]
    assert st.tolist() == g["state"]


@pytest.mark.parametrize("n,a,b", [(21, 0, 1), (24, 22, 23), (30, 3, 29)])
def test_exhaustive_ties_go_to_the_lowest_index(ctx, n, a, b):
    q = np.zeros((n, n))
    q[a, a] = q[b, b] = -1.0
    q[a, b] = 2.0
    st, e = ctx.exhaustive(q)
    assert e == -1.0
    want = np.zeros(n, dtype=np.uint8)
    want[a] = 1  # 2^a < 2^b
    assert (st == want).all()


def test_exhaustive_all_zero_instance(ctx):
    for n in (5, 22):
        st, e = ctx.exhaustive(np.zeros((n, n)))
        assert e == 0.0 and not st.any()


def test_exhaustive_non_finite_coefficients(ctx):
    n = 9
    q = _random_upper(n, 3)
    q[3, 3] = np.nan
    q[1, 5] = np.inf
    q[2, 7] = -np.inf
    st, e = ctx.exhaustive(q)
    ref_st, ref_e = oracle.exhaustive(_sym(q), n, literal=True)
    assert e == ref_e and (st == ref_st).all()
    q = np.zeros((31, 31))
    q[0, 0] = np.nan
    with pytest.raises(onesolver_amd.OneSolverError) as ei:
        ctx.exhaustive(q)
    assert ei.value.code == onesolver_amd.OS_ERR_UNSUPPORTED


def test_exhaustive_rejects_sizes(ctx):
    with pytest.raises(onesolver_amd.OneSolverError) as ei:
        ctx.exhaustive(np.zeros((41, 41)))
    assert ei.value.code == onesolver_amd.OS_ERR_UNSUPPORTED


def test_exhaustive_cli_gpu_device(tmp_path):
    """one-solver-exhaustive --device-type gpu on the config-1 instance."""
    out = tmp_path / "c1.csv"
    r = subprocess.run([onesolver_amd.EXHAUSTIVE_CLI_PATH, "--input",
                        os.path.join(GOLDEN, "c1_dense28.qubo"), "--output", str(out),
                        "--device-type", "gpu"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    with open(os.path.join(GOLDEN, "c1_ground.json")) as f:
        g = json.load(f)
    head, row = out.read_text().splitlines()
    assert head == ",".join(str(i) for i in range(28)) + ",energy"
    vals = row.split(",")
    assert [int(v) for v in vals[:28]] == g["state"]
    assert abs(float(vals[28]) - g["energy"]) < 1e-4  # the CSV keeps 6 significant digits
