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
        onesolver_amd.DenseProblem(ctx, np.zeros((2000, 2000)))
