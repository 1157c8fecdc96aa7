"""Pins the oracle's `.qubo` reader (or_qubo_parse) before it checks the GPU
reader: the reference's own test strings (tests/io_test.cpp), the grammar
cases, the example and KAT files (their flattened matrices against the plain
loader and the reference's flatten expectation), and its numbers against
Python's float(), an independent correctly rounded conversion."""
import glob
import os
import struct

import numpy as np
import pytest

import oracle
import qubo_corpus as qc
from conftest import GOLDEN
from qubo_io import load_qubo


def _flat(res):
    n = res["n"]
    q = np.zeros((n, n))
    for i, j, v in zip(res["i"], res["j"], res["v"]):
        if i == j:
            q[i, i] = v
        else:
            q[i, j] += v
            q[j, i] += v
    return q


def test_reference_io_test_strings():
    code, _, res = oracle.qubo_parse(qc.REF_VALID.encode())
    assert code == 0
    assert res["n"] == 3                                     # io_test.cpp:40-47
    got = {(int(i), int(j)): v for i, j, v in zip(res["i"], res["j"], res["v"])}
    assert got[(0, 0)] == -0.5 and got[(2, 2)] == -0.7       # :49-57
    assert got[(0, 1)] == 2.0 and got[(1, 2)] == 4.0         # :59-67
    for text, code in qc.REF_REJECTED:
        assert oracle.qubo_parse(text.encode())[0] == code, text


@pytest.mark.parametrize("k", range(len(qc.GRAMMAR)))
def test_grammar_cases(k):
    text, want = qc.GRAMMAR[k]
    code, msg, res = oracle.qubo_parse(text.encode("latin-1"))
    assert code == want, (text, msg)
    if code:
        assert msg == oracle.QP_MESSAGES[code]


def test_grammar_case_details():
    # first value of a repeated key wins, counted once
    _, _, r = oracle.qubo_parse(b"p qubo 0 4 1 1\n0 0 1\n0 0 2\n0 1 3\n0 1 4\n")
    assert r["v"].tolist() == [1.0, 3.0] and r["num_linear"] == 1 and r["num_quadratic"] == 1
    # n is the largest index + 1, whatever the header says
    assert oracle.qubo_parse(b"p qubo 0 100 1 1\n7 7 1\n2 40 3\n")[2]["n"] == 41
    # numbers: nan, signed infinities, a nan(...) whose ')' is on the next line
    _, _, r = oracle.qubo_parse(b"p qubo 0 4 1 0\n0 0 -inf\n")
    assert r["v"][0] == -np.inf
    _, _, r = oracle.qubo_parse(b"p qubo 0 4 1 0\n0 0 nan(\n)\n")
    assert np.isnan(r["v"][0])
    _, _, r = oracle.qubo_parse(b"p qubo 0 4 1 0\n0 0 -nan\n")
    assert np.isnan(r["v"][0]) and np.signbit(r["v"][0])


def test_example_and_kat_files():
    files = sorted(glob.glob(os.path.join(GOLDEN, "examples", "*.qubo")) +
                   glob.glob(os.path.join(GOLDEN, "kat", "*.qubo")) +
                   [os.path.join(GOLDEN, "c1_dense28.qubo")])
    assert len(files) >= 12
    for f in files:
        with open(f, "rb") as fh:
            code, msg, res = oracle.qubo_parse(fh.read())
        if os.path.basename(f) == "dwave_doc.qubo":
            assert code == 1, msg  # the reference's grammar rejects this file
            continue
        assert code == 0, (f, msg)
        assert np.array_equal(_flat(res), load_qubo(f)), f


def test_flatten_matches_reference_expectation():
    # tests/qubo_helpers_test.cpp:14-27, the model written as text
    text = (b"p qubo 0 5 4 6\n0 0 0.5\n1 1 -2.0\n2 2 1.0\n4 4 -1.5\n0 1 1.0\n0 3 7.2\n"
            b"1 4 -1.0\n2 3 2.0\n2 4 -1.5\n3 4 -3.5\n")
    code, _, res = oracle.qubo_parse(text)
    assert code == 0 and res["n"] == 5
    expected = [0.5, 1.0, 0.0, 7.2, 0.0, 1.0, -2.0, 0.0, 0.0, -1.0, 0.0, 0.0, 1.0, 2.0, -1.5, 7.2,
                0.0, 2.0, 0.0, -3.5, 0.0, -1.0, -1.5, -3.5, -1.5]
    assert _flat(res).ravel().tolist() == expected


def test_numbers_match_python_float():
    rng = np.random.default_rng(7)
    toks = qc.decimal_tokens(rng, 20000)
    code, msg, res = oracle.qubo_parse(qc.linear_text(toks).encode())
    assert code == 0, msg
    assert len(res["v"]) == len(toks)
    got = res["v"].view(np.uint64)
    want = np.array([struct.unpack("<Q", struct.pack("<d", float(t)))[0] for t in toks],
                    dtype=np.uint64)
    bad = np.nonzero(got != want)[0]
    assert bad.size == 0, [toks[k] for k in bad[:5]]


def test_power_of_five_table_is_current_and_exact():
    """The committed pow5_table.h equals its generator's output, and each
    entry satisfies its definition: T(q) = floor of 5^q normalised to
    [2^127, 2^128), and floor(q log2 10), checked with exact integers."""
    import importlib.util
    import io
    from contextlib import redirect_stdout

    from conftest import ROOT
    csrc = os.path.join(ROOT, "onesolver_amd", "csrc")
    spec = importlib.util.spec_from_file_location("gen_pow5", os.path.join(csrc, "gen_pow5.py"))
    gen = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(gen)
    buf = io.StringIO()
    with redirect_stdout(buf):
        gen.main()
    with open(os.path.join(csrc, "pow5_table.h")) as f:
        assert f.read() == buf.getvalue()
    for q in range(gen.QMIN, gen.QMAX + 1):
        hi, lo, e2 = gen.entry(q)
        t = (hi << 64) | lo
        if q >= 0:
            x, b = 5 ** q, (5 ** q).bit_length()
            # t * 2^(b - 128) <= 5^q < (t + 1) * 2^(b - 128)
            assert (t << max(0, b - 128)) <= (x << max(0, 128 - b)) < ((t + 1) << max(0, b - 128))
        else:
            y = 5 ** (-q)
            b = y.bit_length()
            assert t * y <= (1 << (127 + b)) < (t + 1) * y
        # 2^e2 <= 10^q < 2^(e2 + 1), as rationals
        num, den = (10 ** q, 1) if q >= 0 else (1, 10 ** (-q))
        lo2 = (1 << e2, 1) if e2 >= 0 else (1, 1 << (-e2))
        hi2 = (1 << (e2 + 1), 1) if e2 + 1 >= 0 else (1, 1 << (-(e2 + 1)))
        assert lo2[0] * den <= num * lo2[1] and num * hi2[1] < hi2[0] * den
