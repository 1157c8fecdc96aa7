"""GPU parity of population annealing on the dense path (os_dense_population).

On integer-valued instances the sweeps and the running energies are exact, so
the resampling rounds see the oracle's energies and draw the oracle's
ancestors; states, best energies and ancestor lists then equal the oracle's
(or_population_dense) bit for bit.  The oracle is pinned in
tests/test_oracle.py (equal temperatures reduce to plain sweeps, systematic
resampling's offspring counts, reweighting of Boltzmann samples, a population
anneal ending in the Boltzmann law where plain annealing does not).
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


# (n, R, K, theta, refresh): registers path, a population that is not a
# multiple of 64 (padded rows), FP64 refreshes between rounds, and the
# streaming path (n > 1024)
@pytest.mark.parametrize("n,R,K,theta,refresh", [(100, 64, 6, 4, 0), (300, 200, 5, 3, 6),
                                                 (1000, 128, 4, 2, 0), (1100, 64, 3, 2, 2)])
def test_population_matches_oracle_bit_for_bit(ctx, n, R, K, theta, refresh):
    q = _int_instance(n, n + 1)
    ts = onesolver_amd.tscale_from_beta(np.geomspace(0.2, 3.0, K))
    print(f"[gpu] population n={n} R={R} K={K} theta={theta} refresh={refresh}", flush=True)
    prob = onesolver_amd.DenseProblem(ctx, q)
    out = prob.population(R, ts, theta, seed=9, instance=3, refresh_interval=refresh,
                          final_states=True)
    prob.close()
    fin, bs, be, anc = oracle.population_dense(q, n, ts, theta, R, seed=9, instance=3, threads=8)
    assert (out["ancestors"] == anc).all()
    assert (out["final_states"] == fin).all()
    assert (out["best_energies"] == be).all()
    assert (out["best_states"] == bs).all()
    i = int(np.argmin(be))
    assert out["best_energy"] == be[i] and (out["best_state"] == bs[i]).all()
    # resampling selected: fewer distinct ancestors than replicas in a round
    assert min(len(np.unique(a)) for a in anc) < R


def test_population_batch_instances_use_their_own_streams(ctx):
    n, R, K, theta = 128, 96, 4, 3
    qs = np.stack([_int_instance(n, 50 + i) for i in range(3)])
    ts = onesolver_amd.tscale_from_beta(np.geomspace(0.5, 2.5, K))
    print("[gpu] population batch of 3", flush=True)
    b = onesolver_amd.DenseBatch(ctx, qs)
    out = b.population(R, ts, theta, seed=5, instance=10)
    b.close()
    for i in range(3):
        _, bs, be, anc = oracle.population_dense(qs[i], n, ts, theta, R, seed=5, instance=10 + i)
        assert (out["ancestors"][i] == anc).all()
        assert (out["best_energies"][i] == be).all()
        assert (out["best_states"][i] == bs).all()


def test_population_at_equal_temperatures_is_plain_annealing(ctx):
    n, R, K, theta = 200, 64, 4, 3
    q = _int_instance(n, 77)
    print("[gpu] population at equal temperatures", flush=True)
    prob = onesolver_amd.DenseProblem(ctx, q)
    a = prob.population(R, np.full(K, 0.8, np.float32), theta, seed=4, final_states=True)
    b = prob.anneal(R, np.full(K * theta, 0.8, np.float32), seed=4, final_states=True)
    prob.close()
    assert (a["ancestors"] == np.arange(R)).all()
    assert (a["final_states"] == b["final_states"]).all()
    assert (a["best_energies"] == b["best_energies"]).all()


def test_population_rejects_bad_schedules(ctx):
    prob = onesolver_amd.DenseProblem(ctx, _int_instance(64, 1))
    ts = np.array([2.0, 1.0], np.float32)
    with pytest.raises(onesolver_amd.OneSolverError):
        prob.population(64, ts, 0)                          # no sweeps per step
    with pytest.raises(onesolver_amd.OneSolverError):
        prob.population(64, np.zeros(0, np.float32), 2)     # no steps
    with pytest.raises(onesolver_amd.OneSolverError):
        prob.population(64, ts, 3, refresh_interval=4)      # refresh not a multiple
    with pytest.raises(onesolver_amd.OneSolverError):
        prob.population(64, np.array([2.0, -1.0], np.float32), 2)
    with pytest.raises(onesolver_amd.OneSolverError):
        prob.population(0, ts, 2)                           # empty population
    prob.close()
