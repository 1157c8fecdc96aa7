"""GPU parity of replica exchange on the dense path (os_dense_tempering).

On integer-valued instances the sweeps are exact (bf16 hi + fp32 sums), and
so are the running energies the exchange rounds compare, so the GPU's states,
best energies and final slots equal the oracle's (or_tempering_dense) bit for
bit.  The oracle itself is pinned in tests/test_oracle.py (equal temperatures
reduce to plain sweeps; the exchange decisions; Boltzmann marginals per slot).
"""
import numpy as np
import pytest

import onesolver_amd
import oracle

pytestmark = pytest.mark.gpu


def _int_instance(n, seed, lo=-3, hi=3):
    rng = np.random.default_rng(seed)
    q = np.triu(rng.integers(lo, hi + 1, (n, n)).astype(np.float64))
    return q + np.triu(q, 1).T


@pytest.mark.parametrize("n,K,L,S,iv,refresh", [(100, 8, 16, 40, 4, 0), (300, 4, 16, 12, 3, 6),
                                                 (1000, 16, 4, 6, 2, 0)])
def test_tempering_matches_oracle_bit_for_bit(ctx, n, K, L, S, iv, refresh):
    q = _int_instance(n, n)
    ts = onesolver_amd.tscale_from_beta(np.geomspace(0.3, 3.0, K))
    print(f"[gpu] tempering n={n} K={K} L={L} S={S} interval={iv}", flush=True)
    prob = onesolver_amd.DenseProblem(ctx, q)
    out = prob.tempering(L, ts, S, iv, seed=9, instance=3, refresh_interval=refresh,
                         final_states=True)
    prob.close()
    fin, bs, be, sl = oracle.tempering_dense(q, n, ts, L, S, iv, seed=9, instance=3)
    assert (out["final_slots"] == sl).all()
    assert (out["final_states"] == fin).all()
    assert (out["best_energies"] == be).all()
    assert (out["best_states"] == bs).all()
    i = int(np.argmin(be))
    assert out["best_energy"] == be[i] and (out["best_state"] == bs[i]).all()


def test_tempering_batch_instances_use_their_own_streams(ctx):
    n, K, L, S, iv = 128, 4, 16, 16, 4
    qs = np.stack([_int_instance(n, 40 + i) for i in range(2)])
    ts = onesolver_amd.tscale_from_beta(np.geomspace(0.5, 2.5, K))
    b = onesolver_amd.DenseBatch(ctx, qs)
    out = b.tempering(L, ts, S, iv, seed=5, instance=10)
    b.close()
    for i in range(2):
        _, bs, be, sl = oracle.tempering_dense(qs[i], n, ts, L, S, iv, seed=5, instance=10 + i)
        assert (out["final_slots"][i] == sl).all()
        assert (out["best_energies"][i] == be).all()


def test_tempering_with_equal_temperatures_is_plain_annealing(ctx):
    n, K, L, S = 200, 4, 16, 10
    q = _int_instance(n, 77)
    prob = onesolver_amd.DenseProblem(ctx, q)
    a = prob.tempering(L, np.full(K, 0.8, np.float32), S, 2, seed=4, final_states=True)
    b = prob.anneal(K * L, np.full(S, 0.8, np.float32), seed=4, final_states=True)
    prob.close()
    assert (a["final_states"] == b["final_states"]).all()
    assert (a["best_energies"] == b["best_energies"]).all()


def test_tempering_rejects_bad_shapes(ctx):
    prob = onesolver_amd.DenseProblem(ctx, _int_instance(64, 1))
    ts = np.array([2.0, 1.0], np.float32)
    with pytest.raises(onesolver_amd.OneSolverError):
        prob.tempering(4, ts, 8, 0)                       # no exchange interval
    with pytest.raises(onesolver_amd.OneSolverError):
        prob.tempering(4, ts, 8, 3, refresh_interval=4)   # refresh not a multiple
    with pytest.raises(onesolver_amd.OneSolverError):
        prob.tempering(4, np.array([2.0], np.float32), 8, 2)   # one slot
    with pytest.raises(onesolver_amd.OneSolverError):
        prob.tempering(4, np.array([2.0, -1.0], np.float32), 8, 2)
    prob.close()
    big = onesolver_amd.DenseProblem(ctx, _int_instance(1100, 2))
    with pytest.raises(onesolver_amd.OneSolverError) as ei:
        big.tempering(2, ts, 2, 1)
    assert ei.value.code == onesolver_amd.OS_ERR_INVALID
    big.close()
