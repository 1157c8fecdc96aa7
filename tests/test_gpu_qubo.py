"""GPU parity of the `.qubo` reader (os_qubo_parse) with the oracle's
restatement of parse_qubo + QUBOBuilder (include/model/qubo.hpp:282-417):
the same accept/reject decision and exception message, the same distinct
terms in the same order, the same doubles bit for bit."""
import numpy as np
import pytest

import onesolver_amd
import oracle
import qubo_corpus as qc

pytestmark = pytest.mark.gpu


def _bits(v):
    return np.asarray(v, np.float64).view(np.uint64)


def _check_same(ctx, text, label=""):
    data = text if isinstance(text, bytes) else text.encode("latin-1")
    code, msg, res = oracle.qubo_parse(data)
    if code:
        with pytest.raises(onesolver_amd.InvalidArgument) as ei:
            onesolver_amd.Qubo(ctx, data)
        assert ei.value.message == msg, (label, data[:200])
        return None
    with onesolver_amd.Qubo(ctx, data) as q:
        assert (q.n, q.num_linear, q.num_quadratic) == (res["n"], res["num_linear"],
                                                        res["num_quadratic"]), label
        i, j, v = q.terms()
        assert np.array_equal(i, res["i"]) and np.array_equal(j, res["j"]), label
        assert np.array_equal(_bits(v), _bits(res["v"])), label
    return res


def test_reference_io_test_strings(ctx):
    res = _check_same(ctx, qc.REF_VALID)
    assert res["n"] == 3
    for text, _ in qc.REF_REJECTED:
        _check_same(ctx, text)


def test_grammar_cases(ctx):
    print("[gpu] qubo grammar cases", flush=True)
    for k, (text, _) in enumerate(qc.GRAMMAR):
        _check_same(ctx, text, label=k)


def test_numbers_bit_for_bit(ctx):
    rng = np.random.default_rng(11)
    toks = qc.decimal_tokens(rng, 300000)
    print(f"[gpu] qubo numbers: {len(toks)} tokens", flush=True)
    _check_same(ctx, qc.linear_text(toks))


def test_dense_text_with_duplicates_and_comments(ctx):
    rng = np.random.default_rng(3)
    n = 300
    q = rng.uniform(-1, 1, (n, n))
    lines = qc.dense_text(q).splitlines(keepends=True)
    body = lines[1:]
    dups = [body[k].replace(body[k].split()[2], "%.17g" % rng.uniform(-9, 9))
            for k in rng.integers(0, len(body), 500)]
    body = body + dups
    order = rng.permutation(len(body))
    body = [body[k] for k in order]
    for k in rng.integers(0, len(body), 200):
        body[k] = "c a comment %d\n" % k + body[k]
    res = _check_same(ctx, lines[0] + "".join(body))
    assert res["num_quadratic"] == n * (n - 1) // 2 and res["num_linear"] == n


def test_mixed_line_ends_across_tile_boundaries(ctx):
    rng = np.random.default_rng(5)
    m = 60000
    eols = ["\n", "\r\n", "\r"]
    parts = ["p qubo 0 %d %d 0\n" % (m, m)]
    for k in range(m):
        pad = " " * int(rng.integers(0, 40))
        parts.append("%s%d %d %s%s" % (pad, k, k, "%.*g" % (int(rng.integers(1, 18)),
                                                           rng.normal()),
                                      eols[int(rng.integers(0, 3))]))
    _check_same(ctx, "".join(parts))


def test_sparse_instance_uses_hash_dedupe(ctx):
    rng = np.random.default_rng(9)
    n = 200000
    k = 60000
    i = rng.integers(0, n, k)
    j = rng.integers(0, n, k)
    i, j = np.minimum(i, j), np.maximum(i, j)
    i = np.concatenate([i, i[:3000]])   # repeated keys, later values lose
    j = np.concatenate([j, j[:3000]])
    vals = rng.normal(size=len(i))
    keys = set()
    nl = nq = 0
    for a, b in zip(i, j):
        if (a, b) not in keys:
            keys.add((a, b))
            nl += a == b
            nq += a != b
    text = "p qubo 0 %d %d %d\n" % (n, nl, nq) + "".join(
        "%d %d %.17g\n" % t for t in zip(i, j, vals))
    print(f"[gpu] qubo sparse n={n} lines={len(i)}", flush=True)
    _check_same(ctx, text)


def test_flatten_and_dense_problem(ctx):
    rng = np.random.default_rng(4)
    n = 200
    q = np.triu(rng.integers(-3, 4, (n, n)).astype(np.float64))
    text = qc.dense_text(q)
    res = _check_same(ctx, text)
    flat = np.asarray(oracle.flatten(n, {int(a): v for a, b, v in zip(res["i"], res["j"], res["v"])
                                         if a == b},
                                     {(int(a), int(b)): v for a, b, v in
                                      zip(res["i"], res["j"], res["v"]) if a != b})).reshape(n, n)
    with onesolver_amd.Qubo(ctx, text) as m:
        assert np.array_equal(_bits(m.flatten()), _bits(flat))
        d = m.dense()
    ts = np.geomspace(2.0, 0.2, 20).astype(np.float32)
    a = d.anneal(128, ts, seed=3, final_states=True)
    d.close()
    p = onesolver_amd.DenseProblem(ctx, flat)
    b = p.anneal(128, ts, seed=3, final_states=True)
    p.close()
    assert (a["final_states"] == b["final_states"]).all()
    assert (a["best_energies"] == b["best_energies"]).all()


def test_negative_zero_off_diagonal_flattens_to_plus_zero(ctx):
    with onesolver_amd.Qubo(ctx, "p qubo 0 4 1 1\n0 0 -0\n1 2 -0\n") as m:
        f = m.flatten()
        _, _, v = m.terms()
    assert np.signbit(v).tolist() == [True, True]
    assert np.signbit(f[0, 0]) and not np.signbit(f[1, 2]) and not np.signbit(f[2, 1])


def test_binary_round_trip_and_rejections(ctx):
    rng = np.random.default_rng(8)
    q = rng.uniform(-1, 1, (64, 64))
    text = qc.dense_text(q)
    with onesolver_amd.Qubo(ctx, text) as a:
        blob = a.to_binary()
        ta = a.terms()
        fa = a.flatten()
    assert blob[:4] == b"OSQB" and len(blob) == 32 + 16 * len(ta[0])
    with onesolver_amd.Qubo(ctx, binary=blob) as b:
        tb = b.terms()
        assert (b.n, b.max_nodes) == (64, 64)
        assert all(np.array_equal(x, y) for x, y in zip(ta, tb))
        assert np.array_equal(_bits(b.flatten()), _bits(fa))
    T = len(ta[0])

    def corrupt(off, raw):
        x = bytearray(blob)
        x[off:off + len(raw)] = raw
        return bytes(x)

    bad = [blob[:-1], b"XSQB" + blob[4:], corrupt(4, (2).to_bytes(4, "little")),
           corrupt(32, (5).to_bytes(4, "little")),                       # i[0] > j[0]
           corrupt(32 + 4 * T, (64).to_bytes(4, "little")),              # j[0] = n
           corrupt(8, (65).to_bytes(4, "little")),                       # n != largest + 1
           corrupt(32 + 4 * T + 8, ta[1][1].tobytes()),                  # (0, 2) -> (0, 1)
           corrupt(16, (3).to_bytes(8, "little"))]                       # linear count
    for k, blob_bad in enumerate(bad):
        with pytest.raises(onesolver_amd.OneSolverError):
            onesolver_amd.Qubo(ctx, binary=blob_bad)


def test_index_above_int_range_is_unsupported(ctx):
    with pytest.raises(onesolver_amd.OneSolverError) as ei:
        onesolver_amd.Qubo(ctx, "p qubo 0 4 1 0\n2147483648 2147483648 1\n")
    assert ei.value.code == onesolver_amd.OS_ERR_UNSUPPORTED


def test_large_dense_text(ctx):
    """A 2.1M-line dense text: counts, order and values against the oracle."""
    rng = np.random.default_rng(12)
    n = 2048
    q = rng.uniform(-1, 1, (n, n)).astype(np.float32).astype(np.float64)
    text = qc.dense_text(q, width=5).encode()
    print(f"[gpu] qubo dense n={n}: {len(text) / 1e6:.1f} MB", flush=True)
    res = _check_same(ctx, text)
    assert res["n"] == n and res["num_linear"] == n
    iu = np.triu_indices(n)
    assert np.array_equal(res["v"], q[iu])


def test_device_text_matches_host_text(ctx):
    """os_qubo_parse_device: text already in GPU memory, including texts the
    sequential reader decides and numbers left to strtod (short tokens
    gathered on the GPU, long ones through a host copy)."""
    rng = np.random.default_rng(21)
    toks = qc.decimal_tokens(rng, 3000) + ["0." + "3" * 100, "1" * 70 + "e-60"]
    texts = [t for t, _ in qc.GRAMMAR] + [qc.REF_VALID, qc.linear_text(toks), ""]
    for k, text in enumerate(texts):
        data = text.encode("latin-1")
        code, msg, res = oracle.qubo_parse(data)
        with onesolver_amd.DeviceBytes(ctx, data) as dev:
            if code:
                with pytest.raises(onesolver_amd.InvalidArgument) as ei:
                    onesolver_amd.Qubo(ctx, device=dev.source)
                assert ei.value.message == msg, k
                continue
            with onesolver_amd.Qubo(ctx, device=dev.source) as q:
                i, j, v = q.terms()
                assert q.n == res["n"], k
                assert np.array_equal(i, res["i"]) and np.array_equal(j, res["j"]), k
                assert np.array_equal(_bits(v), _bits(res["v"])), k


def test_long_lines_read_from_global_memory(ctx):
    """Rounds of 256 lines longer than the 16 KiB staging span are parsed from
    global memory; short and long rounds mixed."""
    rng = np.random.default_rng(31)
    m = 5000
    parts = ["p qubo 0 %d %d 0\n" % (m, m)]
    for k in range(m):
        pad = " " * (150 if (k // 256) % 2 else int(rng.integers(0, 4)))
        com = ("c " + "x" * int(rng.integers(0, 300)) + "\n") if rng.random() < 0.1 else ""
        parts.append("%s%s%d %d %.17g\n" % (com, pad, k, k, rng.normal()))
    _check_same(ctx, "".join(parts))


def test_dense_problem_streaming_size_and_file_loading(ctx, tmp_path):
    """os_qubo_dense for n > 1024 (fields in HBM) equals a DenseProblem of the
    flattened matrix; Qubo.load reads text and binary files alike."""
    rng = np.random.default_rng(14)
    n = 1100
    q = np.triu(rng.integers(-2, 3, (n, n)).astype(np.float64))
    text = qc.dense_text(q)
    (tmp_path / "a.qubo").write_text(text)
    with onesolver_amd.Qubo.load(ctx, tmp_path / "a.qubo") as m:
        (tmp_path / "a.osqb").write_bytes(m.to_binary())
        flat = m.flatten()
    assert np.array_equal(flat, q + np.triu(q, 1).T)
    ts = np.geomspace(2.0, 0.5, 3).astype(np.float32)
    with onesolver_amd.Qubo.load(ctx, tmp_path / "a.osqb") as m:
        d = m.dense()
    a = d.anneal(64, ts, seed=2, final_states=True)
    d.close()
    p = onesolver_amd.DenseProblem(ctx, flat)
    b = p.anneal(64, ts, seed=2, final_states=True)
    p.close()
    assert (a["final_states"] == b["final_states"]).all()
    assert (a["best_energies"] == b["best_energies"]).all()
