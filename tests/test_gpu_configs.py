"""GPU runs of the benchmark configurations at their real sizes.

Dense SK N=1024 (16384 replicas) and dense N=8192 (4096 replicas) take the
same paths as the benchmark: the register-resident kernel for n <= 1024 and
the streaming path with fields in HBM above.  Each run is checked against the
oracle: bit for bit where the couplings are integers (bf16 hi + lo and the
fp32 field sums are exact), and otherwise through the FP64 energy function of
every returned state (1e-9 relative), the field drift against the FP64
refresh, and the law of the best energies on the same random streams.
"""
import os

import numpy as np
import pytest

import onesolver_amd
import oracle

pytestmark = pytest.mark.gpu

THREADS = min(16, os.cpu_count() or 1)


def _tscale(S, bmin=0.1, bmax=3.0):
    return onesolver_amd.tscale_from_beta(np.geomspace(bmin, bmax, S))


def _int_instance(n, seed, lo=-3, hi=3):
    rng = np.random.default_rng(seed)
    q = np.triu(rng.integers(lo, hi + 1, (n, n), dtype=np.int8)).astype(np.float64)
    q += np.triu(q, 1).T
    return q


def test_c2_sk1024_best_energies_follow_the_oracle(ctx):
    """SK N=1024 generated on the GPU (the benchmark's instance): 256
    replicas, 16 sweeps of the geometric schedule 0.1 -> 3.0.  Same streams
    as or_sweep_dense; the bf16 split moves single fields by ~1e-5, so the
    chains may part, but the best energies keep the oracle's law."""
    n, R, S = 1024, 256, 16
    print(f"[gpu] C2 SK n={n} R={R} S={S} vs oracle", flush=True)
    b = onesolver_amd.DenseBatch(ctx, sk=(1, n, 0, 0))
    q = b.q()[0]
    ts = _tscale(S)
    out = b.anneal(R, ts, seed=11, instance=0, refresh_interval=8)
    b.close()
    e = out["best_energies"][0]
    ref_e = oracle.energy(q, out["best_states"][0])
    assert np.all(np.abs(e - ref_e) <= 1e-9 * np.abs(ref_e))
    _, _, be_or = oracle.sweep_dense(q, n, ts, R, seed=11, instance=0, threads=THREADS)
    d = e - be_or
    print(f"[gpu] C2 gpu mean {e.mean():.4f} oracle mean {be_or.mean():.4f} "
          f"identical {np.mean(d == 0):.3f}", flush=True)
    assert abs(d.mean()) <= 4 * d.std() / np.sqrt(R) + 1e-9
    # most chains have not parted after 16 sweeps
    assert np.mean(np.abs(d) <= 1e-6 * np.abs(be_or)) > 0.5


def test_c2_sk1024_all_replicas_energies_and_drift(ctx):
    """The full C2 replica count: 16384 replicas, 64 sweeps, FP64 refresh
    every 32.  Every returned energy equals the FP64 energy function of its
    state within 1e-9 relative; the fp32 field drift stays below 1e-3."""
    n, R, S = 1024, 16384, 64
    print(f"[gpu] C2 SK n={n} R={R} S={S} energies", flush=True)
    b = onesolver_amd.DenseBatch(ctx, sk=(1, n, 0, 0))
    q = b.q()[0]
    out = b.anneal(R, _tscale(S), seed=1, refresh_interval=32)
    b.close()
    e = out["best_energies"][0]
    ref_e = oracle.energy(q, out["best_states"][0])
    assert np.all(np.abs(e - ref_e) <= 1e-9 * np.abs(ref_e))
    drift = out["stats"]["max_field_drift"]
    print(f"[gpu] C2 best {e.min():.4f} drift {drift:.2e}", flush=True)
    assert 0.0 < drift < 1e-3
    i = int(np.argmin(e))
    assert out["best_energy"][0] == e[i] and (out["best_state"][0] == out["best_states"][0][i]).all()
    # SK ground energy ~ -0.763 n (Parisi); 64 sweeps come within a few percent
    assert e.min() < -0.65 * n


def test_c5_integer_8192_matches_oracle_bit_for_bit(ctx):
    """N=8192 (fields in HBM, 64-spin blocks, field passes beside the
    decisions) on an integer instance: 64 replicas, 2 sweeps, equal to
    or_sweep_dense state for state."""
    n, R, S = 8192, 64, 2
    print(f"[gpu] C5 integer n={n} R={R} S={S}", flush=True)
    q = _int_instance(n, 8192)
    ts = _tscale(S)
    prob = onesolver_amd.DenseProblem(ctx, q)
    out = prob.anneal(R, ts, seed=4, instance=1, refresh_interval=1, final_states=True)
    prob.close()
    fin, bs, be = oracle.sweep_dense(q, n, ts, R, seed=4, instance=1, threads=THREADS)
    assert (out["final_states"] == fin).all()
    assert (out["best_states"] == bs).all()
    np.testing.assert_allclose(out["best_energies"], oracle.energy(q, bs), rtol=1e-12)
    assert out["stats"]["max_field_drift"] == 0.0


def test_c5_uniform_8192_all_replicas_energies_and_drift(ctx):
    """The C5 instance (dense uniform N=8192 generated on the GPU) with the
    full 4096 replicas for 2 sweeps, FP64 refresh after each: every returned
    energy equals the FP64 energy function within 1e-9 relative, and the
    drift of the fp32 fields stays below 1e-2 (fields of magnitude ~50)."""
    n, R, S = 8192, 4096, 2
    print(f"[gpu] C5 uniform n={n} R={R} S={S} energies", flush=True)
    b = onesolver_amd.DenseBatch(ctx, uniform=(1, n, 0, 0))
    q = b.q()[0]
    out = b.anneal(R, _tscale(S), seed=1, refresh_interval=1)
    b.close()
    e = out["best_energies"][0]
    ref_e = oracle.energy(q, out["best_states"][0])
    assert np.all(np.abs(e - ref_e) <= 1e-9 * np.abs(ref_e))
    drift = out["stats"]["max_field_drift"]
    print(f"[gpu] C5 best {e.min():.3f} drift {drift:.2e}", flush=True)
    assert 0.0 < drift < 1e-2
    assert out["stats"]["pass_launches"] > 0
