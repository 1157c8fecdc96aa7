"""GPU parity of the dense sweep-mode annealer (bf16 hi+lo MFMA delayed updates).

Integer-valued couplings are exact in bf16 hi and in fp32 sums, so on those
the GPU chains must equal the oracle's (or_sweep_dense) bit for bit.  On
real-valued instances the fields carry the bf16 split error; there the
returned energies are checked against the FP64 energy function, the field
drift is bounded, and the annealed energies are compared statistically.
"""
import numpy as np
import pytest

import onesolver_amd
import oracle
from onesolver_amd import instances

pytestmark = pytest.mark.gpu


def _int_instance(n, seed, lo=-3, hi=3):
    rng = np.random.default_rng(seed)
    q = rng.integers(lo, hi + 1, (n, n)).astype(np.float64)
    q = np.triu(q)
    return q + np.triu(q, 1).T


def _tscale(S, bmin=0.1, bmax=3.0):
    return onesolver_amd.tscale_from_beta(np.geomspace(bmin, bmax, S))


@pytest.mark.parametrize("n,R,S,refresh", [(100, 128, 40, 0), (300, 64, 12, 5), (1000, 64, 6, 0),
                                           (128, 192, 30, 7)])
def test_dense_integer_instances_match_oracle_bit_for_bit(ctx, n, R, S, refresh):
    q = _int_instance(n, n)
    ts = _tscale(S)
    print(f"[gpu] dense exact n={n} R={R} S={S}", flush=True)
    prob = onesolver_amd.DenseProblem(ctx, q)
    out = prob.anneal(R, ts, seed=7, instance=3, refresh_interval=refresh, final_states=True)
    fin, bs, be = oracle.sweep_dense(q, n, ts, R, seed=7, instance=3)
    assert (out["final_states"] == fin).all()
    assert (out["best_states"] == bs).all()
    assert np.array_equal(out["best_energies"], be)
    assert out["stats"]["max_field_drift"] == 0.0
    i = int(np.argmin(be))
    assert out["best_energy"] == be[i] and (out["best_state"] == bs[i]).all()


def test_dense_sk_real_valued(ctx):
    n, R, S = 256, 512, 200
    q32, _ = instances.sk_qubo(n, seed=1, instance=0)
    q = q32.astype(np.float64)
    ts = _tscale(S)
    prob = onesolver_amd.DenseProblem(ctx, q)
    out = prob.anneal(R, ts, seed=11, refresh_interval=50)
    ref_e = oracle.energy(q, out["best_states"])
    assert np.all(np.abs(out["best_energies"] - ref_e) <= 1e-9 * np.abs(ref_e))
    assert out["stats"]["max_field_drift"] < 1e-3
    _, _, be_or = oracle.sweep_dense(q, n, ts, R, seed=11)
    d = out["best_energies"] - be_or
    # same streams, fields differ by the bf16 split: same law of the best energy
    assert abs(d.mean()) < 4 * d.std() / np.sqrt(R) + 1e-9
    print(f"[gpu] SK n={n}: gpu mean {out['best_energies'].mean():.4f} "
          f"oracle mean {be_or.mean():.4f} drift {out['stats']['max_field_drift']:.2e}", flush=True)


def test_dense_greedy_limit_reaches_local_minima(ctx):
    n, R = 200, 64
    q = instances.dense_uniform(n, seed=3).astype(np.float64)
    prob = onesolver_amd.DenseProblem(ctx, q)
    out = prob.anneal(R, np.zeros(30, np.float32), seed=5, final_states=True)
    for x in out["final_states"][:8]:
        h = np.diag(q) + (q - np.diag(np.diag(q))) @ x
        de = np.where(x == 1, -h, h)
        assert (de >= -1e-9).all()


def test_dense_start_states_and_marginals(ctx):
    n, R = 64, 64
    q = _int_instance(n, 2)
    prob = onesolver_amd.DenseProblem(ctx, q)
    init = np.random.default_rng(0).integers(0, 2, (R, n)).astype(np.uint8)
    out = prob.anneal(R, np.zeros(0, np.float32), init_states=init, final_states=True)
    assert (out["final_states"] == init).all() and (out["best_states"] == init).all()
    assert np.array_equal(out["best_energies"], oracle.energy(q, init))
    p = np.where(np.arange(n) % 3 == 0, 1.0, 0.0).astype(np.float32)
    out = prob.anneal(R, np.zeros(0, np.float32), init_marginals=p, final_states=True)
    assert (out["final_states"] == (p > 0.5)).all()
    p = np.full(n, 0.3, np.float32)
    out = prob.anneal(R, np.zeros(0, np.float32), seed=4, init_marginals=p, final_states=True)
    assert (out["final_states"] == oracle.warm_init(p, R, seed=4)).all()
    out = prob.anneal(R, np.zeros(0, np.float32), seed=4, final_states=True)
    assert (out["final_states"] == oracle.sweep_init(n, R, seed=4)).all()


def test_dense_rejects_oversized(ctx):
    with pytest.raises(onesolver_amd.OneSolverError):
        onesolver_amd.DenseBatch(ctx, sk=(1, 20000, 0, 0))


@pytest.mark.parametrize("n,R,S,refresh", [(1100, 64, 3, 0), (2048, 100, 2, 1)])
def test_dense_streaming_path_matches_oracle_bit_for_bit(ctx, n, R, S, refresh):
    """n > 1024: fields in HBM, blocks of 64 spins, one pass per block."""
    q = _int_instance(n, n + 1)
    ts = _tscale(S)
    print(f"[gpu] dense streaming n={n} R={R} S={S}", flush=True)
    prob = onesolver_amd.DenseProblem(ctx, q)
    out = prob.anneal(R, ts, seed=9, instance=2, refresh_interval=refresh, final_states=True)
    fin, bs, be = oracle.sweep_dense(q, n, ts, R, seed=9, instance=2)
    assert (out["final_states"] == fin).all()
    assert (out["best_states"] == bs).all()
    np.testing.assert_allclose(out["best_energies"], oracle.energy(q, bs), rtol=1e-12)
    assert out["stats"]["max_field_drift"] == 0.0


@pytest.mark.parametrize("group,refresh", [(1, 0), (3, 1), (7, 0)])
def test_dense_streaming_pass_grouping(ctx, monkeypatch, group, refresh):
    """Any number of blocks per field pass (OS_STREAM_GROUP), with partial
    last groups and segment ends, gives the oracle's chains, for a batch of
    two instances."""
    n, R, S = 1100, 64, 3
    qs = np.stack([_int_instance(n, 5), _int_instance(n, 6)])
    ts = _tscale(S)
    monkeypatch.setenv("OS_STREAM_GROUP", str(group))
    print(f"[gpu] dense streaming n={n} group={group} refresh={refresh}", flush=True)
    b = onesolver_amd.DenseBatch(ctx, qs)
    out = b.anneal(R, ts, seed=3, instance=4, refresh_interval=refresh, final_states=True)
    b.close()
    nb = -(-n // 64)
    segs = 1 if refresh == 0 else -(-S // refresh)
    per_seg = S * nb if refresh == 0 else refresh * nb
    assert out["stats"]["pass_launches"] == segs * -(-per_seg // group)
    for i in range(2):
        fin, bs, be = oracle.sweep_dense(qs[i], n, ts, R, seed=3, instance=4 + i)
        assert (out["final_states"][i] == fin).all()
        assert (out["best_states"][i] == bs).all()
        np.testing.assert_allclose(out["best_energies"][i], oracle.energy(qs[i], bs), rtol=1e-12)


@pytest.mark.parametrize("depth,group", [(2, 1), (2, 5), (3, 1), (3, 4), (4, 3)])
def test_dense_streaming_pipeline_depth(ctx, monkeypatch, depth, group):
    """Pipeline depth P (OS_STREAM_DEPTH): the decides of group H read the
    fields after pass H-P plus up to P*group - 1 pending blocks, from P field
    copies.  Every depth gives the oracle's chains on a two-instance integer
    batch, and the same bits on a real-valued SK instance (every field is the
    same fp32 sum in the same block order, whichever launch added a block)."""
    n, R, S = 1100, 64, 3
    qs = np.stack([_int_instance(n, 15), _int_instance(n, 16)])
    ts = _tscale(S)
    monkeypatch.setenv("OS_STREAM_GROUP", str(group))
    monkeypatch.setenv("OS_STREAM_DEPTH", str(depth))
    print(f"[gpu] dense streaming depth={depth} group={group}", flush=True)
    b = onesolver_amd.DenseBatch(ctx, qs)
    out = b.anneal(R, ts, seed=2, instance=6, final_states=True)
    b.close()
    for i in range(2):
        fin, bs, be = oracle.sweep_dense(qs[i], n, ts, R, seed=2, instance=6 + i)
        assert (out["final_states"][i] == fin).all()
        assert (out["best_states"][i] == bs).all()
    sk = onesolver_amd.DenseBatch(ctx, sk=(1, 1536, 4, 0))
    got = sk.anneal(R, ts, seed=8, final_states=True)
    monkeypatch.setenv("OS_STREAM_DEPTH", "2")
    monkeypatch.setenv("OS_STREAM_GROUP", "1")
    ref = sk.anneal(R, ts, seed=8, final_states=True)
    sk.close()
    for k in ("final_states", "best_states", "best_energies"):
        assert (got[k] == ref[k]).all(), k


@pytest.mark.parametrize("variant", [20, 21, 30, 31, 40, 50, 60, 61, 62])
def test_dense_streaming_decide_chunking(ctx, monkeypatch, variant):
    """Every operand chunking of the decide launches (OS_STREAM_DECIDE_VARIANT:
    blocks per round trip, next chunk requested early or not, 6x: one operand
    load shared by both replica halves of a column half) gives the oracle's
    chains; 7 blocks per pass leave up to 13 pending blocks at the default
    depth of 2."""
    n, R, S = 1100, 64, 3
    q = _int_instance(n, 8)
    ts = _tscale(S)
    monkeypatch.setenv("OS_STREAM_GROUP", "7")
    monkeypatch.setenv("OS_STREAM_DECIDE_VARIANT", str(variant))
    print(f"[gpu] dense streaming decide variant {variant}", flush=True)
    prob = onesolver_amd.DenseProblem(ctx, q)
    out = prob.anneal(R, ts, seed=6, instance=1, final_states=True)
    prob.close()
    fin, bs, be = oracle.sweep_dense(q, n, ts, R, seed=6, instance=1)
    assert (out["final_states"] == fin).all()
    assert (out["best_states"] == bs).all()


def test_dense_streaming_autotuner(ctx, monkeypatch):
    """os_dense_tune picks blocks per field pass and pass work-groups per CU
    from probe runs and keeps them on the handle.  Neither changes a result
    (every field is the same fp32 sum in the same block order), so runs after
    tuning equal the runs before it on a real-valued instance, and the oracle's
    chains on an integer one.  For n <= 1024 nothing is tuned."""
    for k in ("OS_STREAM_GROUP", "OS_STREAM_PASS_WGS_PER_CU", "OS_STREAM_PASS_GRID"):
        monkeypatch.delenv(k, raising=False)
    n, R, S = 2048, 64, 3
    ts = _tscale(S)
    print("[gpu] dense streaming autotuner, SK n=2048", flush=True)
    b = onesolver_amd.DenseBatch(ctx, sk=(1, n, 9, 0))
    before = b.anneal(R, ts, seed=5, final_states=True)
    t = b.tune(R, ts, seed=5)
    assert t["group"] in (3, 4, 5, 6, 8)
    assert t["pass_wgs_per_cu"] in (1, 2)
    assert t["probe_ms"] > 0.0
    after = b.anneal(R, ts, seed=5, final_states=True)
    assert after["stats"]["pass_launches"] == -(-(S * (n // 64)) // t["group"])
    for k in ("final_states", "best_states", "best_energies"):
        assert (after[k] == before[k]).all(), k
    b.close()
    print("[gpu] dense streaming autotuner, integer n=1100", flush=True)
    q = _int_instance(1100, 12)
    prob = onesolver_amd.DenseProblem(ctx, q)
    prob.tune(R, ts, seed=7)
    out = prob.anneal(R, ts, seed=7, final_states=True)
    prob.close()
    fin, bs, be = oracle.sweep_dense(q, 1100, ts, R, seed=7, instance=0)
    assert (out["final_states"] == fin).all()
    assert (out["best_states"] == bs).all()
    small = onesolver_amd.DenseProblem(ctx, _int_instance(100, 1))
    assert small.tune(R, ts) == {"group": 0, "pass_wgs_per_cu": 0, "probe_ms": 0.0}
    small.close()


def test_dense_streaming_sk_energies_and_drift(ctx):
    n, R, S = 2048, 64, 6
    b = onesolver_amd.DenseBatch(ctx, sk=(1, n, 4, 0))
    q = b.q()[0]
    ts = _tscale(S)
    print("[gpu] dense streaming SK n=2048", flush=True)
    out = b.anneal(R, ts, seed=2, refresh_interval=3)
    e64 = oracle.energy(q, out["best_states"][0])
    np.testing.assert_allclose(out["best_energies"][0], e64, rtol=1e-9)
    assert 0.0 < out["stats"]["max_field_drift"] < 1e-3
    # a few sweeps from random states already go far below the random-state energy
    assert out["best_energy"][0] < -0.3 * n
    b.close()


def test_dense_batch_equals_single_instance_runs(ctx):
    """A batch of instances in one launch gives every instance exactly the
    chains of a run on its own (stream id instance + i) and of the oracle."""
    qs = np.stack([_int_instance(100, s) for s in (1, 2, 3)])
    R, S = 96, 15
    ts = _tscale(S)
    print("[gpu] dense batch of 3, n=100", flush=True)
    batch = onesolver_amd.DenseBatch(ctx, qs)
    out = batch.anneal(R, ts, seed=5, instance=10, refresh_interval=4, final_states=True)
    for i in range(3):
        fin, bs, be = oracle.sweep_dense(qs[i], 100, ts, R, seed=5, instance=10 + i)
        assert (out["final_states"][i] == fin).all()
        assert (out["best_states"][i] == bs).all()
        np.testing.assert_allclose(out["best_energies"][i], oracle.energy(qs[i], bs), rtol=1e-12)
        single = onesolver_amd.DenseProblem(ctx, qs[i]).anneal(R, ts, seed=5, instance=10 + i,
                                                                refresh_interval=4)
        assert (single["best_states"] == out["best_states"][i]).all()
        assert single["best_energy"] == out["best_energy"][i]
    # warm start per instance
    p = np.stack([np.full(100, 0.2, np.float32), np.full(100, 0.5, np.float32),
                  np.full(100, 0.9, np.float32)])
    w = batch.anneal(R, ts[:0], seed=5, instance=10, init_marginals=p)
    for i in range(3):
        ref = oracle.warm_init(p[i], R, seed=5, instance=10 + i)
        assert (w["best_states"][i] == ref).all()
    batch.close()


def test_sk_instances_generated_on_the_gpu_match_the_host_generator(ctx):
    """The GPU generator restates instances.sk_qubo: same Philox counters and
    Box-Muller transform; off-diagonal couplings agree to fp32 rounding (the
    device and host FP64 log/cos may differ in the last bit), diagonals to the
    summation order."""
    I, n = 3, 96
    print("[gpu] SK generator n=96", flush=True)
    b = onesolver_amd.DenseBatch(ctx, sk=(I, n, 7, 20))
    qg = b.q()
    for s in range(I):
        qh, _ = instances.sk_qubo(n, seed=7, instance=20 + s)
        off = ~np.eye(n, dtype=bool)
        np.testing.assert_allclose(qg[s][off], qh[off], rtol=2 ** -22, atol=0)
        np.testing.assert_allclose(np.diag(qg[s]), np.diag(qh), rtol=1e-6, atol=1e-7)
        assert (qg[s] == qg[s].T).all()
    # the packed operands of a generated batch anneal like an uploaded one
    ts = _tscale(8)
    out = b.anneal(64, ts, seed=3, instance=20)
    up = onesolver_amd.DenseBatch(ctx, qg).anneal(64, ts, seed=3, instance=20)
    assert (out["best_states"] == up["best_states"]).all()
    b.close()


def test_uniform_instances_generated_on_the_gpu_match_the_host_generator(ctx):
    """Bit for bit: the uniform generator uses only exact fp32 arithmetic."""
    b = onesolver_amd.DenseBatch(ctx, uniform=(2, 70, 0, 5))
    qg = b.q()
    for s in range(2):
        assert (qg[s] == instances.dense_uniform(70, seed=0, instance=5 + s)).all()
    b.close()
