"""Best-of-k band check of reference-mode annealing against the reference's
published runs.

benchmarks/annealing/results/<n>/001.csv holds one row per run of
`one-solver-anneal --num-iter I --num-tries K --schedule-type S --beta-max 10`
(benchmarks/annealing/Snakefile:54-66).  The Snakefile passes no --beta-min,
so every run used the CLI default 0.1 (app/one-solver-anneal.cpp:66-67) and
the beta_min column is only a label.  Each row is the lowest energy over K
replicas of I random-site attempts (annealing.hpp:97-125), beta used as a
temperature and rising from 0.1 to 10 (linear: to 10.1).

A sampler returns the per-replica best energies of `count` independent
replicas for a schedule.  For a published point (schedule, I, K) the blocks of
K consecutive replicas give the sampler's law of the best-of-K energy, and the
point's PIT is the fraction of blocks below the published value (ties count
one half).  The reference's law and the sampler's agree when the published
values sit inside the sampler's band at every point.
"""
import numpy as np

BETA_MIN = 0.1
BETA_MAX = 10.0
# the published energies carry 6 significant digits (include/model/solution.hpp:58-67)
TIE = 1e-3


def schedule(oracle, kind, num_iter, beta_min=BETA_MIN, beta_max=BETA_MAX):
    if kind == "linear":
        return oracle.linear_schedule(beta_min, beta_max, num_iter)
    return oracle.geometric_schedule(beta_min, beta_max, num_iter)


def published_points(grid):
    """{(schedule, num_iter, num_tries): energy} over the grid's rows (the
    rows that differ only in the beta_min label are identical)."""
    return {(r["schedule"], r["num_iter"], r["num_tries"]): r["energy"] for r in grid}


def pit(samples, e):
    s = np.asarray(samples)
    return float(((s < e - TIE).sum() + 0.5 * (np.abs(s - e) <= TIE).sum()) / len(s))


def point_pits(sampler, points, cells, ks, blocks):
    """PIT of every published (schedule, num_iter, k) for the cells
    (schedule, num_iter) and best-of-k sizes ks, from `blocks` blocks of
    max(ks) replicas per cell."""
    kmax = max(ks)
    out = {}
    for kind, it in cells:
        be = np.asarray(sampler(kind, it, blocks * kmax)).reshape(blocks, kmax)
        run_min = np.minimum.accumulate(be, axis=1)
        for k in ks:
            out[(kind, it, k)] = pit(run_min[:, k - 1], points[(kind, it, k)])
    return out


def band_verdict(pits, lo=0.005, hi=0.995, mean_lo=0.25, mean_hi=0.85):
    """(ok, points outside [lo, hi], mean PIT).  A sampler whose best-of-k
    is much better (PIT near 1) or worse (near 0) than the reference's fails
    at single points; a smaller shift of the whole law moves the mean."""
    outside = {key: p for key, p in pits.items() if not lo <= p <= hi}
    mean = float(np.mean(list(pits.values())))
    return (not outside and mean_lo <= mean <= mean_hi), outside, mean
