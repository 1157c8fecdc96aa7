"""Pins the CPU oracle before it is trusted as the checker of the GPU path.

Each test checks an oracle function against something independent of it:
published known-answer vectors, the reference's own test data and published
outputs, closed forms, libm, brute force, or a second restatement.
"""
import json
import os
import subprocess

import numpy as np
import pytest

import oracle
from conftest import GOLDEN
from qubo_io import load_qubo, published_grid

# --------------------------------------------------------------------------
# Philox4x32-10
# --------------------------------------------------------------------------

RANDOM123_KAT = [
    # (counter, key, output) -- Random123 kat_vectors, philox4x32_10
    ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
    ((0xFFFFFFFF,) * 4, (0xFFFFFFFF, 0xFFFFFFFF), (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
    ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
     (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
]


@pytest.mark.parametrize("ctr,key,out", RANDOM123_KAT)
def test_philox_known_answers(ctr, key, out):
    assert tuple(int(v) for v in oracle.philox(ctr, key)) == out


def test_philox_matches_rocrand_header(tmp_path):
    src = os.path.join(os.path.dirname(GOLDEN), "rocrand_kat.cpp")
    exe = tmp_path / "kat"
    r = subprocess.run(["g++", "-O1", "-std=c++17", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
                        src, "-o", str(exe)], capture_output=True, text=True)
    if r.returncode != 0:
        pytest.skip("rocRAND host header not compilable here: " + r.stderr[-300:])
    cases = [(1234, 0), (1234, 4), (0, 8), (0xDEADBEEFCAFEF00D, 4 * 1000), (7, 4 * 123456)]
    inp = "".join(f"{s} {o}\n" for s, o in cases)
    out = subprocess.run([str(exe)], input=inp, capture_output=True, text=True, check=True).stdout
    for (seed, off), line in zip(cases, out.strip().splitlines()):
        got = oracle.philox((off // 4, 0, 0, 0), (seed & 0xFFFFFFFF, seed >> 32))
        assert [int(v) for v in line.split()] == [int(v) for v in got]
    # the survey's first words for seed 1234
    assert [f"{int(v):08x}" for v in oracle.philox((0, 0, 0, 0), (1234, 0))] == \
        ["2090b348", "da7cf0ab", "4401906f", "cbca470e"]


def test_neg_ln_u_against_libm_on_whole_grid():
    k = np.arange(1 << 23, dtype=np.uint64)
    w = (k << np.uint64(9)).astype(np.uint32)
    got = oracle.neg_ln_u_many(w).astype(np.float64)
    u = (2.0 * k.astype(np.float64) + 1.0) * 2.0 ** -24
    ref = -np.log(u)
    rel = np.abs(got - ref) / ref
    assert rel.max() < 4e-7
    assert (got > 0).all()
    # the low 9 bits of the word are ignored
    assert oracle.neg_ln_u(0x12345600) == oracle.neg_ln_u(0x123457FF)


# --------------------------------------------------------------------------
# flatten, energy, schedules
# --------------------------------------------------------------------------


def test_flatten_reference_kat():
    # tests/qubo_helpers_test.cpp:14-27
    lin = {0: 0.5, 1: -2.0, 2: 1.0, 4: -1.5}
    quad = {(0, 1): 1.0, (0, 3): 7.2, (1, 4): -1.0, (2, 3): 2.0, (2, 4): -1.5, (3, 4): -3.5}
    expected = [0.5, 1.0, 0.0, 7.2, 0.0, 1.0, -2.0, 0.0, 0.0, -1.0, 0.0, 0.0, 1.0, 2.0, -1.5, 7.2,
                0.0, 2.0, 0.0, -3.5, 0.0, -1.0, -1.5, -3.5, -1.5]
    assert list(oracle.flatten(5, lin, quad)) == expected


def test_energy_closed_form_random():
    rng = np.random.default_rng(3)
    for n in (1, 2, 7, 31):
        q = rng.uniform(-1, 1, (n, n))
        q = (q + q.T) / 2
        for _ in range(5):
            x = rng.integers(0, 2, n).astype(np.uint8)
            ref = sum(q[i, j] * x[i] * x[j] for i in range(n) for j in range(i, n))
            assert oracle.energy(q, x) == pytest.approx(ref, rel=1e-14, abs=1e-14)


def _kats():
    with open(os.path.join(GOLDEN, "kat", "expected.json")) as f:
        return [(os.path.join(GOLDEN, "kat", e["file"]), e["energy"]) for e in json.load(f)]


@pytest.mark.parametrize("path,energy", _kats())
def test_exhaustive_reference_kats(path, energy):
    # tests/exhaustive_test.cpp:146-157, tolerance 1e-13 (:156)
    q = load_qubo(path)
    n = q.shape[0]
    s1, e1 = oracle.exhaustive(q, n, literal=True)
    s2, e2 = oracle.exhaustive(q, n, literal=False)
    assert abs(e1 - energy) < 1e-13
    assert abs(e2 - energy) < 1e-13
    assert (s1 == s2).all()
    assert abs(oracle.energy(q, s1) - energy) < 1e-13


EXAMPLES = {"simple": -2.0, "test1": -12.0, "test2": -1.2, "csp5": -22.0, "csp7": -14.0,
            "csp13": -32.0}


@pytest.mark.parametrize("name", sorted(EXAMPLES))
def test_example_instances_ground_energies(name):
    q = load_qubo(os.path.join(GOLDEN, "examples", name + ".qubo"))
    _, e = oracle.exhaustive(q, q.shape[0])
    assert abs(e - EXAMPLES[name]) < 1e-12


def test_exhaustive_literal_vs_gray_random():
    rng = np.random.default_rng(11)
    for n in (3, 9, 14):
        for _ in range(3):
            q = rng.uniform(-1, 1, (n, n))
            q = (q + q.T) / 2
            s1, e1 = oracle.exhaustive(q, n, literal=True)
            s2, e2 = oracle.exhaustive(q, n, literal=False, threads=4)
            assert e1 == e2 and (s1 == s2).all()


def test_exhaustive_ties_go_to_lowest_index():
    # all-zero model: every state has energy 0, the reference returns state 0
    s, e = oracle.exhaustive(np.zeros((6, 6)), 6)
    assert e == 0.0 and not s.any()
    # x0 and x1 symmetric: minimum at {0} or {1}; index 1 (x0 = 1) is lower
    q = np.array([[-1.0, 5.0], [5.0, -1.0]])
    s, e = oracle.exhaustive(q, 2, literal=True)
    assert list(s) == [1, 0] and e == -1.0


def _published_rows(n):
    """(state, printed energy) of every row of results/<n>/001.csv."""
    for row in published_grid(n):
        yield row["state"], row["energy"]


@pytest.mark.parametrize("n", [128, 512])
def test_energy_matches_published_annealer_outputs(n):
    # benchmarks/annealing/results/<n>/001.csv: state + energy printed with 6
    # significant digits (include/model/solution.hpp:58-67)
    q = load_qubo(os.path.join(GOLDEN, "chimera", f"{n}_001.qubo"))
    count = 0
    for x, printed in _published_rows(n):
        e = oracle.energy(q, x)
        tol = 0.5 * 10 ** (np.floor(np.log10(abs(e))) - 5)
        assert abs(e - printed) <= tol * 1.0001
        count += 1
    assert count == 1200


@pytest.mark.parametrize("n,expected", [(128, -235.866654), (512, -928.746690),
                                        (2048, -3418.266732)])
def test_tensor_network_ground_states(n, expected):
    q = load_qubo(os.path.join(GOLDEN, "chimera", f"{n}_001.qubo"))
    with open(os.path.join(GOLDEN, "chimera", f"{n}_groundstate_TN.txt")) as f:
        tok = f.read().split()
    spins = np.array([int(v) for v in tok[3:3 + n]])
    x = ((spins + 1) // 2).astype(np.uint8)
    e = oracle.energy(q, x)
    assert abs(e - expected) < 1e-5
    if n in (128, 512):  # the published annealer never beat the ground state
        best_pub = min(p for _, p in _published_rows(n))
        assert e < best_pub


def test_schedules_closed_forms():
    lin = oracle.linear_schedule(0.1, 10.0, 1000)
    i = np.arange(1000)
    assert np.array_equal(lin, 0.1 + 10.0 * i / 999.0)
    assert lin[-1] == 0.1 + 10.0  # ends at beta_min + beta_max
    assert np.isnan(oracle.linear_schedule(0.1, 1.0, 1)[0])
    geo = oracle.geometric_schedule(0.1, 3.0, 4096)
    alpha = (3.0 / 0.1) ** (1.0 / 4095)
    ref = [0.1]
    for _ in range(4095):
        ref.append(ref[-1] * alpha)
    assert np.array_equal(geo, np.array(ref))
    assert abs(geo[-1] - 3.0) < 1e-12


# --------------------------------------------------------------------------
# reference-mode annealer
# --------------------------------------------------------------------------


def test_anneal_literal_and_fields_follow_the_same_chain_on_integer_instances():
    for path, ground in _kats():
        q = load_qubo(path)
        n = q.shape[0]
        beta = oracle.geometric_schedule(0.1, 10.0, 300)
        s1, e1 = oracle.anneal(q, n, beta, 64, literal=True)
        s2, e2 = oracle.anneal(q, n, beta, 64, literal=False)
        assert (s1 == s2).all()
        if np.array_equal(q, np.round(q)):  # exact arithmetic: identical energies
            assert (e1 == e2).all()
        assert np.allclose(e1, e2, rtol=0, atol=1e-12)
        assert (e1 >= ground - 1e-12).all()
        assert np.allclose(oracle.energy(q, s1), e1, rtol=0, atol=1e-12)


def test_anneal_literal_vs_fields_random_real_instance():
    q = load_qubo(os.path.join(GOLDEN, "c1_dense28.qubo"))
    beta = oracle.geometric_schedule(0.1, 1.0, 1000)
    s1, e1 = oracle.anneal(q, 28, beta, 64, literal=True)
    s2, e2 = oracle.anneal(q, 28, beta, 64, literal=False)
    assert (s1 == s2).mean() > 0.999
    assert np.allclose(e1, e2, rtol=1e-12, atol=1e-12)


def test_anneal_replicas_are_independent_of_num_tries():
    q = load_qubo(os.path.join(GOLDEN, "kat", "kat_4.qubo"))
    n = q.shape[0]
    beta = oracle.linear_schedule(0.1, 2.0, 200)
    s_all, e_all = oracle.anneal(q, n, beta, 40)
    s_sub, e_sub = oracle.anneal(q, n, beta, 10, first_replica=30)
    assert (s_all[30:] == s_sub).all() and (e_all[30:] == e_sub).all()


def test_anneal_heats_up_under_reference_semantics():
    # beta is a temperature (annealing.hpp:107): at beta -> 0 only downhill
    # moves and ties... are rejected/accepted as the reference does; a hot
    # (large beta) schedule accepts almost everything.
    q = load_qubo(os.path.join(GOLDEN, "kat", "kat_2.qubo"))
    n = q.shape[0]
    cold = np.full(400, 1e-9)
    s, e = oracle.anneal(q, n, cold, 16)
    for x in s:  # a cold chain ends in a local minimum of single flips
        for j in range(n):
            y = x.copy()
            y[j] ^= 1
            assert oracle.energy(q, y) >= oracle.energy(q, x) - 1e-12


# --------------------------------------------------------------------------
# sweep-mode annealers
# --------------------------------------------------------------------------


def _random_csr(n, d, seed, integer_diag=False):
    from onesolver_amd import instances

    rp, col, val, diag = instances.regular_csr(n, d, seed=seed, instance=1)
    if integer_diag:
        diag = (np.round(diag * 2) / 2).astype(np.float32)
    return rp, col, val, diag


def test_sweep_greedy_limit_ends_in_local_minima():
    q = load_qubo(os.path.join(GOLDEN, "c1_dense28.qubo"))
    fin, best, be = oracle.sweep_dense(q, 28, np.zeros(20, np.float32), 32)
    for x in fin:
        e = oracle.energy(q, x)
        for j in range(28):
            y = x.copy()
            y[j] ^= 1
            assert oracle.energy(q, y) >= e - 1e-12
    assert np.allclose(oracle.energy(q, best), be, atol=1e-9)


def test_sweep_csr_equals_sweep_dense_on_exact_instances():
    from onesolver_amd import instances

    rp, col, val, diag = _random_csr(60, 4, 5, integer_diag=True)
    q = instances.csr_to_dense(rp, col, val, diag)
    ts = (1.0 / np.geomspace(0.1, 3.0, 50)).astype(np.float32)
    f1, b1, e1 = oracle.sweep_dense(q, 60, ts, 16, seed=9)
    f2, b2, e2 = oracle.sweep_csr(rp, col, val, diag, ts, 16, seed=9)
    assert (f1 == f2).all() and (b1 == b2).all() and (e1 == e2).all()
    assert np.allclose(oracle.csr_energy(rp, col, val, diag, b2), e2, atol=1e-9)


@pytest.mark.parametrize("kind", ["dense", "csr"])
def test_sweep_stationary_law_is_boltzmann(kind):
    # Sequential-sweep Metropolis leaves the Boltzmann law invariant: after
    # many sweeps at fixed beta the final states of independent replicas
    # follow p(x) ~ exp(-beta E(x)).
    q = np.array([[0.3, -1.0, 0.5, 0.0], [-1.0, -0.2, 0.7, 0.4], [0.5, 0.7, 0.1, -0.6],
                  [0.0, 0.4, -0.6, 0.2]])
    beta = 1.3
    R = 100000
    ts = np.full(30, 1.0 / beta, dtype=np.float32)
    if kind == "dense":
        fin, _, _ = oracle.sweep_dense(q, 4, ts, R, seed=21)
    else:
        rows, cols = np.nonzero(q - np.diag(np.diag(q)))
        rp = np.searchsorted(rows, np.arange(5)).astype(np.int32)
        fin, _, _ = oracle.sweep_csr(rp, cols.astype(np.int32), q[rows, cols].astype(np.float32),
                                     np.diag(q).astype(np.float32), ts, R, seed=21)
    idx = fin @ (1 << np.arange(4))
    counts = np.bincount(idx, minlength=16)
    states = ((np.arange(16)[:, None] >> np.arange(4)) & 1).astype(np.uint8)
    e = oracle.energy(q.astype(np.float32).astype(np.float64), states)
    p = np.exp(-beta * (e - e.min()))
    p /= p.sum()
    chi2 = ((counts - R * p) ** 2 / (R * p)).sum()
    assert chi2 < 45.0  # 15 dof: p ~ 1e-4


def test_warm_init_follows_marginals():
    p = np.array([0.0, 1.0, 0.25, 0.9], dtype=np.float32)
    st = oracle.warm_init(p, 20000, seed=3)
    m = st.mean(axis=0)
    assert m[0] == 0.0 and m[1] == 1.0
    assert abs(m[2] - 0.25) < 0.015 and abs(m[3] - 0.9) < 0.01


def test_tempering_with_equal_temperatures_is_plain_sweeps():
    rng = np.random.default_rng(4)
    n = 40
    q = rng.integers(-3, 4, (n, n)).astype(np.float64)
    q = np.triu(q) + np.triu(q, 1).T
    K, L, S = 4, 8, 24
    fin, bs, be, sl = oracle.tempering_dense(q, n, np.full(K, 1.5, np.float32), L, S, 4, seed=3,
                                             instance=5)
    fin2, bs2, be2 = oracle.sweep_dense(q, n, np.full(S, 1.5, np.float32), L * K, seed=3,
                                        instance=5)
    assert (fin == fin2).all() and (bs == bs2).all() and (be == be2).all()
    for l in range(L):
        assert sorted(sl[l * K:(l + 1) * K]) == list(range(K))


def test_tempering_exchange_decisions():
    # one exchange round after one sweep between a hot (beta 0.2) and a cold
    # (beta 1) slot: with d = (beta_0 - beta_1)(E_0 - E_1), a pair must swap
    # when d >= 0, keeps its slots when d < -15, and in between swaps with
    # probability exp(d)
    rng = np.random.default_rng(8)
    n = 24
    q = rng.integers(-4, 5, (n, n)).astype(np.float64)
    q = np.triu(q) + np.triu(q, 1).T
    ts = np.array([5.0, 1.0], np.float32)
    L, iv = 512, 1
    _, _, _, sl = oracle.tempering_dense(q, n, ts, L, 2 * iv, iv, seed=11, instance=2)
    # the states at the exchange: plain sweeps of each slot's replicas
    hot, _, _ = oracle.sweep_dense(q, n, np.full(iv, ts[0]), 2 * L, seed=11, instance=2)
    cold, _, _ = oracle.sweep_dense(q, n, np.full(iv, ts[1]), 2 * L, seed=11, instance=2)
    e_hot = oracle.energy(q, hot[0::2])
    e_cold = oracle.energy(q, cold[1::2])
    d = (1.0 / ts[0] - 1.0 / ts[1]) * (e_hot - e_cold)
    swapped = sl[0::2] == 1
    assert (swapped[d >= 0]).all() and (d >= 0).sum() > 50
    assert not (swapped[d < -15]).any() and (d < -15).sum() > 50
    mid = (d < 0) & (d >= -15)
    p = np.exp(d[mid])
    assert abs(swapped[mid].sum() - p.sum()) < 4 * np.sqrt((p * (1 - p)).sum()) + 1


def test_tempering_keeps_boltzmann_marginals():
    # replica exchange leaves each slot's states distributed as
    # exp(-beta_k E(x)) at stationarity
    q = np.array([[0.3, -1.0, 0.5, 0.0], [-1.0, -0.2, 0.7, 0.4], [0.5, 0.7, 0.1, -0.6],
                  [0.0, 0.4, -0.6, 0.2]])
    betas = np.array([0.4, 1.3])
    ts = (1.0 / betas).astype(np.float32)
    L = 60000
    fin, _, _, sl = oracle.tempering_dense(q, 4, ts, L, 30, 1, seed=21)
    states = ((np.arange(16)[:, None] >> np.arange(4)) & 1).astype(np.uint8)
    e = oracle.energy(q.astype(np.float32).astype(np.float64), states)
    for k, beta in enumerate(betas):
        idx = fin[sl == k] @ (1 << np.arange(4))
        counts = np.bincount(idx, minlength=16)
        p = np.exp(-beta * (e - e.min()))
        p /= p.sum()
        chi2 = ((counts - L * p) ** 2 / (L * p)).sum()
        assert chi2 < 45.0, (k, chi2)  # 15 dof: p ~ 1e-4


# --------------------------------------------------------------------------
# population annealing
# --------------------------------------------------------------------------

_PA_Q = 2.0 * np.array([[0.3, -1.0, 0.5, 0.0], [-1.0, -0.2, 0.7, 0.4], [0.5, 0.7, 0.1, -0.6],
                        [0.0, 0.4, -0.6, 0.2]])
_PA_STATES = ((np.arange(16)[:, None] >> np.arange(4)) & 1).astype(np.uint8)


def _boltzmann_chi2(states, q, beta):
    """Pearson chi^2 (15 dof) of 4-spin states against exp(-beta E)."""
    e = oracle.energy(q.astype(np.float32).astype(np.float64), _PA_STATES)
    counts = np.bincount(states @ (1 << np.arange(4)), minlength=16)
    p = np.exp(-beta * (e - e.min()))
    p /= p.sum()
    R = len(states)
    return ((counts - R * p) ** 2 / (R * p)).sum()


def test_resample_at_equal_weights_keeps_every_replica():
    rng = np.random.default_rng(1)
    E = rng.normal(size=300)
    # (u = 0 exactly puts the targets on the cumulative weights, where
    # rounding of (j + u) / R * W decides the tie)
    for u in (1e-6, 0.37, 0.999999):
        assert (oracle.resample(E, 0.0, u) == np.arange(300)).all()
        assert (oracle.resample(np.full(300, 2.5), 1.7, u) == np.arange(300)).all()


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_resample_is_systematic(seed):
    # systematic resampling: ancestors ascend, replica i gets floor or ceil of
    # R p_i offspring, and over u the mean count is R p_i exactly, where p_i is
    # the importance weight exp(-dbeta E_i) normalised
    rng = np.random.default_rng(seed)
    R = 500
    E = rng.normal(scale=3.0, size=R)
    dbeta = 0.6
    p = np.exp(-dbeta * (E - E.min()))
    p /= p.sum()
    mean = np.zeros(R)
    M = 2000
    for m in range(M):
        anc = oracle.resample(E, dbeta, (m + 0.5) / M)
        assert (np.diff(anc.astype(np.int64)) >= 0).all()
        c = np.bincount(anc, minlength=R)
        assert c.sum() == R
        assert ((c >= np.floor(R * p - 1e-9)) & (c <= np.ceil(R * p + 1e-9))).all()
        mean += c
    assert np.abs(mean / M - R * p).max() < 2.0 / M * 2 + 1e-9


@pytest.mark.parametrize("b0,b1", [(0.5, 1.0), (1.0, 2.0), (0.2, 1.5)])
def test_resample_moves_boltzmann_samples_to_the_colder_temperature(b0, b1):
    # i.i.d. samples of exp(-b0 E) resampled with dbeta = b1 - b0 follow
    # exp(-b1 E); the same samples not resampled do not
    e = oracle.energy(_PA_Q.astype(np.float32).astype(np.float64), _PA_STATES)
    p0 = np.exp(-b0 * (e - e.min()))
    p0 /= p0.sum()
    rng = np.random.default_rng(int(10 * b0 + b1))
    idx = rng.choice(16, size=60000, p=p0)
    anc = oracle.resample(e[idx], b1 - b0, rng.random())
    # chi^2 inflated by the weights' effective sample size: 70 for 15 dof
    assert _boltzmann_chi2(_PA_STATES[idx[anc]], _PA_Q, b1) < 70.0
    assert _boltzmann_chi2(_PA_STATES[idx], _PA_Q, b1) > 1000.0


def test_resample_uniform_is_a_53_bit_uniform():
    u = np.array([oracle.resample_uniform(7, 3, k) for k in range(20000)])
    assert ((u >= 0.0) & (u < 1.0)).all()
    assert (u * 2.0 ** 53 == np.floor(u * 2.0 ** 53)).all()
    assert (u * 2.0 ** 29 != np.floor(u * 2.0 ** 29)).any()  # more than 32 bits
    assert abs(u.mean() - 0.5) < 0.01
    assert np.abs(np.sort(u) - (np.arange(20000) + 0.5) / 20000).max() < 0.015  # KS
    assert oracle.resample_uniform(7, 4, 0) != oracle.resample_uniform(7, 3, 0)


def test_population_at_equal_temperatures_is_plain_sweeps():
    rng = np.random.default_rng(6)
    n = 40
    q = rng.integers(-3, 4, (n, n)).astype(np.float64)
    q = np.triu(q) + np.triu(q, 1).T
    K, theta, R = 5, 3, 48
    fin, bs, be, anc = oracle.population_dense(q, n, np.full(K, 1.2, np.float32), theta, R,
                                               seed=3, instance=5)
    fin2, bs2, be2 = oracle.sweep_dense(q, n, np.full(K * theta, 1.2, np.float32), R, seed=3,
                                        instance=5)
    assert (fin == fin2).all() and (bs == bs2).all() and (be == be2).all()
    assert anc.shape == (K - 1, R) and (anc == np.arange(R)).all()


def test_population_anneal_ends_in_boltzmann():
    # one sweep per temperature from beta 0.1 to 3: the resampled population
    # ends distributed as exp(-3 E); plain annealing on the same schedule
    # (no resampling) is far from it
    betas = np.geomspace(0.1, 3.0, 20)
    ts = (1.0 / betas).astype(np.float32)
    R = 60000
    fin, _, _, anc = oracle.population_dense(_PA_Q, 4, ts, 1, R, seed=1, threads=8)
    assert _boltzmann_chi2(fin, _PA_Q, 3.0) < 45.0  # 15 dof: p ~ 1e-4
    assert len(np.unique(anc[-1])) < R  # resampling did select
    sa, _, _ = oracle.sweep_dense(_PA_Q, 4, ts, R, seed=1, threads=8)
    assert _boltzmann_chi2(sa, _PA_Q, 3.0) > 1000.0


@pytest.mark.parametrize("coloured", [False, True])
def test_population_csr_at_equal_temperatures_is_plain_sweeps(coloured):
    # half-integer diagonal and +-1 weights: energies exact, so the FP64
    # recompute at every step start equals the running energy
    rp, col, val, diag = _random_csr(80, 5, 7, integer_diag=True)
    order = oracle.csr_colour_order(rp, col)[0] if coloured else None
    K, theta, R = 4, 3, 40
    fin, bs, be, anc = oracle.population_csr(rp, col, val, diag, np.full(K, 0.9, np.float32),
                                             theta, R, seed=3, instance=2, order=order)
    fin2, bs2, be2 = oracle.sweep_csr(rp, col, val, diag, np.full(K * theta, 0.9, np.float32), R,
                                      seed=3, instance=2, order=order)
    assert (anc == np.arange(R)).all()
    assert (fin == fin2).all() and (bs == bs2).all() and (be == be2).all()


def test_population_csr_anneal_ends_in_boltzmann():
    # the 4-spin instance as CSR: the sparse population ends distributed as
    # exp(-3 E); plain sparse sweeps on the same schedule do not
    q = _PA_Q
    rp, col, val, diag = _pa_q_csr()
    betas = np.geomspace(0.1, 3.0, 20)
    ts = (1.0 / betas).astype(np.float32)
    R = 60000
    fin, _, _, anc = oracle.population_csr(rp, col, val, diag, ts, 1, R, seed=2, threads=8)
    assert _boltzmann_chi2(fin, q, 3.0) < 45.0  # 15 dof: p ~ 1e-4
    assert len(np.unique(anc[-1])) < R
    sa, _, _ = oracle.sweep_csr(rp, col, val, diag, ts, R, seed=2, threads=8)
    assert _boltzmann_chi2(sa, q, 3.0) > 1000.0


def test_population_csr_best_is_the_lowest_over_the_steps():
    # best energies are attained by the best states and never above the
    # step-0 best (a plain run of the first step's sweeps)
    rp, col, val, diag = _random_csr(120, 4, 9)
    ts = (1.0 / np.geomspace(0.3, 3.0, 6)).astype(np.float32)
    R, theta = 64, 4
    fin, bs, be, anc = oracle.population_csr(rp, col, val, diag, ts, theta, R, seed=5)
    e = oracle.csr_energy(rp, col, val, diag, bs)
    assert np.allclose(e, be, rtol=0, atol=1e-5)  # running sums of fp32 dE
    _, _, be0 = oracle.sweep_csr(rp, col, val, diag, np.full(theta, ts[0], np.float32), R, seed=5)
    assert (be <= be0).all() and (be < be0).any()
    assert anc.shape == (5, R) and (np.diff(anc.astype(np.int64), axis=1) >= 0).all()


def _pa_q_csr():
    """The 4-spin population-test instance as CSR (both (j,l) and (l,j))."""
    rp, col, val = [0], [], []
    for j in range(4):
        for l in range(4):
            if l != j:
                col.append(l)
                val.append(_PA_Q[j, l])
        rp.append(len(col))
    return rp, col, val, np.diag(_PA_Q)


@pytest.mark.parametrize("coloured", [False, True])
def test_tempering_csr_with_equal_temperatures_is_plain_sweeps(coloured):
    rp, col, val, diag = _random_csr(80, 5, 11, integer_diag=True)
    order = oracle.csr_colour_order(rp, col)[0] if coloured else None
    K, L, S = 4, 8, 14
    fin, bs, be, sl = oracle.tempering_csr(rp, col, val, diag, np.full(K, 1.3, np.float32), L, S,
                                           4, seed=3, instance=5, order=order)
    fin2, bs2, be2 = oracle.sweep_csr(rp, col, val, diag, np.full(S, 1.3, np.float32), L * K,
                                      seed=3, instance=5, order=order)
    assert (fin == fin2).all() and (bs == bs2).all() and (be == be2).all()
    for l in range(L):
        assert sorted(sl[l * K:(l + 1) * K]) == list(range(K))


def test_tempering_csr_exchange_decisions():
    # one exchange round after one sweep between a hot and a cold slot: pairs
    # with d >= 0 swap, pairs with d < -15 keep their slots, and in between
    # the swap count matches sum exp(d)
    rp, col, val, diag = _random_csr(40, 4, 12, integer_diag=True)
    ts = np.array([5.0, 1.0], np.float32)
    L = 512
    _, _, _, sl = oracle.tempering_csr(rp, col, val, diag, ts, L, 2, 1, seed=11, instance=2)
    hot, _, _ = oracle.sweep_csr(rp, col, val, diag, np.full(1, ts[0]), 2 * L, seed=11,
                                 instance=2)
    cold, _, _ = oracle.sweep_csr(rp, col, val, diag, np.full(1, ts[1]), 2 * L, seed=11,
                                  instance=2)
    e_hot = oracle.csr_energy(rp, col, val, diag, hot[0::2])
    e_cold = oracle.csr_energy(rp, col, val, diag, cold[1::2])
    d = (1.0 / ts[0] - 1.0 / ts[1]) * (e_hot - e_cold)
    swapped = sl[0::2] == 1
    assert (swapped[d >= 0]).all() and (d >= 0).sum() > 10
    assert not (swapped[d < -12]).any() and (d < -12).sum() > 10
    mid = (d < 0) & (d >= -12)
    p = np.exp(d[mid])
    assert mid.sum() > 300
    assert abs(swapped[mid].sum() - p.sum()) < 4 * np.sqrt((p * (1 - p)).sum()) + 1


def test_tempering_csr_keeps_boltzmann_marginals():
    rp, col, val, diag = _pa_q_csr()
    betas = np.array([0.4, 1.3])
    ts = (1.0 / betas).astype(np.float32)
    L = 60000
    fin, _, _, sl = oracle.tempering_csr(rp, col, val, diag, ts, L, 30, 1, seed=21, threads=8)
    for k, beta in enumerate(betas):
        assert _boltzmann_chi2(fin[sl == k], _PA_Q, beta) < 45.0, k  # 15 dof: p ~ 1e-4


# --------------------------------------------------------------------------
# sparse sweeps in colour-class order
# --------------------------------------------------------------------------


def _greedy_colours(rp, col):
    """Greedy colouring in index order, restated: smallest colour not held by
    a lower-index neighbour."""
    n = len(rp) - 1
    colour = np.zeros(n, dtype=np.int64)
    for j in range(n):
        taken = {int(colour[l]) for l in col[rp[j]:rp[j + 1]] if l < j}
        c = 0
        while c in taken:
            c += 1
        colour[j] = c
    return colour


@pytest.mark.parametrize("n,d,seed", [(60, 4, 1), (257, 6, 2), (1000, 3, 3), (9, 8, 4)])
def test_csr_colour_order_is_the_greedy_colouring(n, d, seed):
    rp, col, val, diag = _random_csr(n, d, seed)
    order, cs = oracle.csr_colour_order(rp, col)
    colour = _greedy_colours(rp, col)
    assert sorted(order.tolist()) == list(range(n))
    assert cs[0] == 0 and cs[-1] == n and len(cs) == colour.max() + 2
    for c in range(len(cs) - 1):
        cls = order[cs[c]:cs[c + 1]]
        assert (colour[cls] == c).all() and (np.diff(cls) > 0).all()
    rows = np.repeat(np.arange(n), np.diff(rp))
    assert (colour[rows] != colour[col]).all()  # no coupling inside a class


def test_sweep_csr_in_identity_order_is_the_sequential_sweep():
    rp, col, val, diag = _random_csr(80, 5, 6)
    ts = (1.0 / np.geomspace(0.1, 3.0, 12)).astype(np.float32)
    a = oracle.sweep_csr(rp, col, val, diag, ts, 16, seed=3)
    b = oracle.sweep_csr(rp, col, val, diag, ts, 16, seed=3, order=np.arange(80, dtype=np.int32))
    for x, y in zip(a, b):
        assert (x == y).all()
    # a different order is a different chain with the same thresholds per site
    o, _ = oracle.csr_colour_order(rp, col)
    c = oracle.sweep_csr(rp, col, val, diag, ts, 16, seed=3, order=o)
    assert not (a[0] == c[0]).all()


def test_sweep_csr_colour_order_greedy_limit_and_boltzmann_law():
    # tscale 0: only downhill flips, so the final states are single-flip minima
    from onesolver_amd import instances

    rp, col, val, diag = _random_csr(120, 4, 8)
    q = instances.csr_to_dense(rp, col, val, diag)
    o, _ = oracle.csr_colour_order(rp, col)
    fin, _, _ = oracle.sweep_csr(rp, col, val, diag, np.zeros(30, np.float32), 24, seed=2, order=o)
    for x in fin:
        h = np.diag(q) + (q - np.diag(np.diag(q))) @ x
        assert (np.where(x == 1, -h, h) >= -1e-6).all()
    # at fixed beta the class-ordered sweep keeps exp(-beta E) invariant
    q4 = np.array([[0.3, -1.0, 0.5, 0.0], [-1.0, -0.2, 0.7, 0.4], [0.5, 0.7, 0.1, -0.6],
                   [0.0, 0.4, -0.6, 0.2]])
    rows, cols = np.nonzero(q4 - np.diag(np.diag(q4)))
    rp4 = np.searchsorted(rows, np.arange(5)).astype(np.int32)
    o4, cs4 = oracle.csr_colour_order(rp4, cols.astype(np.int32))
    assert not (o4 == np.arange(4)).all()  # the visiting order really changes
    beta, R = 1.3, 100000
    fin, _, _ = oracle.sweep_csr(rp4, cols.astype(np.int32), q4[rows, cols].astype(np.float32),
                                 np.diag(q4).astype(np.float32),
                                 np.full(30, 1.0 / beta, np.float32), R, seed=21, order=o4)
    counts = np.bincount(fin @ (1 << np.arange(4)), minlength=16)
    states = ((np.arange(16)[:, None] >> np.arange(4)) & 1).astype(np.uint8)
    e = oracle.energy(q4.astype(np.float32).astype(np.float64), states)
    p = np.exp(-beta * (e - e.min()))
    p /= p.sum()
    chi2 = ((counts - R * p) ** 2 / (R * p)).sum()
    assert chi2 < 45.0  # 15 dof: p ~ 1e-4
