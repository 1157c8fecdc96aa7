"""Pins the reference-mode annealer's law against the reference's own runs.

The oracle's reference-mode chain (or_anneal_fields, and or_anneal_literal,
which follows it state for state) restates sa::anneal
(include/simulated_annealing/annealing.hpp:55-140).  Its trajectories cannot
be compared with the reference's (oneMKL's distributions are not available
here), so its law is compared with the reference's published runs on the
Chimera instances: benchmarks/annealing/results/{128,512}/001.csv, made by
benchmarks/annealing/Snakefile:54-66.  A second check fixes the acceptance law
at finite beta against the Boltzmann distribution exp(-E/beta) on a 4-spin
instance.  Each check is shown to fail for planted mistakes: one sweep of N
attempts per iteration instead of one attempt (Q1), beta doubled, and beta
used as an inverse temperature.
"""
import os

import numpy as np
import pytest

import oracle
import published_band as pb
from conftest import GOLDEN
from qubo_io import load_qubo, published_grid

THREADS = min(8, os.cpu_count() or 1)


def _instance(n):
    return load_qubo(os.path.join(GOLDEN, "chimera", f"{n}_001.qubo"))


def _fields_sampler(q, n, seed, spb=1, beta_scale=1.0):
    """Independent replicas (or_anneal_fields): the law the GPU follows."""
    def sample(kind, it, count):
        beta = pb.schedule(oracle, kind, it) * beta_scale
        _, be = oracle.anneal(q, n, beta, count, sweeps_per_beta=spb, seed=seed, threads=THREADS)
        return be
    return sample


def _stream_sampler(q, n, seed, spb=1, beta_scale=1.0, block=460):
    """Runs of `block` replicas that share one stream at offsets 0, 1, 2, ...
    (or_anneal_stream), as the reference's replicas do (annealing.hpp:87)."""
    def sample(kind, it, count):
        beta = pb.schedule(oracle, kind, it) * beta_scale
        _, be = oracle.anneal_shared_stream(q, n, beta, count, sweeps_per_beta=spb, seed=seed,
                                            threads=THREADS, block_tries=block)
        return be
    return sample


@pytest.mark.parametrize("n", [128, 512])
def test_published_grid_is_one_schedule_and_nested_in_num_tries(n):
    """The grid's structure pins two facts about the reference's runs that
    the band check relies on: beta_min was never passed (rows differing only
    in the beta_min label are identical, so every run was deterministic and
    used beta 0.1 -> 10), and a run of K tries holds the replicas of every
    smaller run (replica i's engine depends on i only, annealing.hpp:87), so
    the best energy never rises with K."""
    grid = published_grid(n)
    assert len(grid) == 1200
    by_cell = {}
    for r in grid:
        key = (r["schedule"], r["num_iter"], r["num_tries"])
        by_cell.setdefault(key, []).append(r)
    assert len(by_cell) == 200
    for rows in by_cell.values():
        assert sorted(r["beta_min"] for r in rows) == [0.1, 0.3, 0.5, 0.7, 0.9, 1.1]
        assert all(r["energy"] == rows[0]["energy"] for r in rows)
        assert all((r["state"] == rows[0]["state"]).all() for r in rows)
    pts = pb.published_points(grid)
    for kind in ("linear", "geometric"):
        for it in range(100, 1100, 100):
            curve = [pts[(kind, it, k)] for k in range(10, 500, 50)]
            assert all(b <= a for a, b in zip(curve, curve[1:]))


BAND_CASES = {
    # n: (cells, blocks of 460 replicas)
    128: ([(s, it) for s in ("linear", "geometric") for it in (100, 200, 500, 1000)], 200),
    512: ([(s, it) for s in ("linear", "geometric") for it in (100, 400, 1000)], 100),
}


def _report(n, pits, mean):
    print(f"n={n} mean PIT {mean:.3f}; " + ", ".join(f"{k[0][0]}{k[1]}/{k[2]}:{p:.2f}"
                                                    for k, p in sorted(pits.items())))


@pytest.mark.parametrize("n", [128, 512])
def test_published_runs_fall_inside_the_oracle_band(n):
    """Every published best-of-K energy (K = 10 and 460 tries) lies inside the
    0.5-99.5% band of the oracle's best-of-K energies, and the mean PIT is
    within [0.25, 0.85].  The sampler shares one stream among each run's
    replicas as the reference does (replica i's engine starts i outputs into
    it, annealing.hpp:87), so its best-of-K carries the same coupling of
    neighbouring replicas."""
    q = _instance(n)
    cells, blocks = BAND_CASES[n]
    pts = pb.published_points(published_grid(n))
    pits = pb.point_pits(_stream_sampler(q, n, seed=20 + n), pts, cells, (10, 460), blocks)
    ok, outside, mean = pb.band_verdict(pits)
    _report(n, pits, mean)
    assert ok, (outside, mean)


@pytest.mark.parametrize("n", [128, 512])
def test_published_runs_against_independent_replicas(n):
    """Best-of-10 on independent streams (or_anneal_fields, the law of the GPU
    kernel) holds every published K = 10 point inside its 0.5-99.5% band.
    Ten neighbouring replicas share few draws; at K = 460 the published runs
    are a little worse than independent replicas would give (N=512, linear
    schedule, 1000 iterations: above all of 100 independent best-of-460
    values, at PIT 0.9 for the shared-stream sampler)."""
    q = _instance(n)
    cells, _ = BAND_CASES[n]
    pts = pb.published_points(published_grid(n))
    pits = pb.point_pits(_fields_sampler(q, n, seed=40 + n), pts, cells, (10,), 1000)
    ok, outside, mean = pb.band_verdict(pits)
    _report(n, pits, mean)
    assert ok, (outside, mean)


def test_band_rejects_one_sweep_per_iteration():
    """Planted Q1 mistake: N attempts per schedule step (a sweep) instead of
    one.  The oracle then anneals far better than the reference did, and
    every published point falls above its band."""
    n = 128
    q = _instance(n)
    cells = [("linear", 100), ("geometric", 100)]
    pts = pb.published_points(published_grid(n))
    pits = pb.point_pits(_stream_sampler(q, n, seed=7, spb=n), pts, cells, (10, 460), 20)
    ok, outside, mean = pb.band_verdict(pits)
    assert not ok and len(outside) == len(pits) and mean == 1.0
    pits = pb.point_pits(_fields_sampler(q, n, seed=7, spb=n), pts, cells, (10,), 100)
    assert not pb.band_verdict(pits)[0]


@pytest.mark.parametrize("scale", [2.0, 0.5])
def test_band_rejects_a_wrong_temperature_scale(scale):
    """Planted mistakes in the acceptance law's temperature: beta doubled,
    or halved, in the same check as the published runs pass."""
    n = 128
    q = _instance(n)
    cells, blocks = BAND_CASES[n]
    pts = pb.published_points(published_grid(n))
    pits = pb.point_pits(_stream_sampler(q, n, seed=20 + n, beta_scale=scale), pts, cells,
                         (10, 460), blocks)
    ok, outside, mean = pb.band_verdict(pits)
    _report(n, pits, mean)
    assert not ok, (pits, mean)


# --------------------------------------------------------------------------
# finite-beta law: Boltzmann distribution of the chain at a fixed beta
# --------------------------------------------------------------------------

Q4 = np.array([[-1.0, 2.0, -0.5, 0.0],
               [2.0, 0.5, 1.0, -1.5],
               [-0.5, 1.0, -0.75, 0.5],
               [0.0, -1.5, 0.5, 0.25]])


def _boltzmann_chi2(beta_run, beta_law, R=20000, attempts=400, seed=3):
    """Chi-square of the final states of R reference-mode chains run at the
    constant beta_run against exp(-E/beta_law) over the 16 states."""
    n = 4
    _, _, fin = oracle.anneal(Q4, n, np.full(attempts, beta_run), R, seed=seed, threads=THREADS,
                              final=True)
    idx = fin.astype(np.int64) @ (1 << np.arange(n))
    counts = np.bincount(idx, minlength=16).astype(np.float64)
    states = ((np.arange(16)[:, None] >> np.arange(n)) & 1).astype(np.uint8)
    e = oracle.energy(Q4, states)
    p = np.exp(-(e - e.min()) / beta_law)
    p /= p.sum()
    return float(((counts - R * p) ** 2 / (R * p)).sum())


@pytest.mark.parametrize("beta", [0.7, 2.0])
def test_reference_chain_is_boltzmann_with_beta_as_temperature(beta):
    """Random-site Metropolis with accept iff E_new < E_cur or
    exp((E_cur - E_new)/beta) > u (annealing.hpp:106-107) has the stationary
    law exp(-E/beta): beta is a temperature.  400 attempts (100 per spin) from
    uniform random bits reach it on 4 spins."""
    chi2 = _boltzmann_chi2(beta, beta)
    assert chi2 < 45.0, chi2  # 15 degrees of freedom: p ~ 1e-4


@pytest.mark.parametrize("beta_run", [4.0, 0.5])
def test_boltzmann_check_rejects_wrong_temperature_laws(beta_run):
    """Planted mistakes in the oracle's chain, checked against the law
    exp(-E/2): the chain run at beta doubled (4.0), and the chain that uses
    beta as an inverse temperature, exp(-beta dE), which is the reference
    chain at temperature 1/beta (0.5).  Both fail."""
    assert _boltzmann_chi2(beta_run, 2.0) > 200.0
