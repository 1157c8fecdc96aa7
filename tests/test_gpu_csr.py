"""GPU parity of the batched sparse (CSR) annealer.

The sweep kernel computes each local field as diag_j followed by the set
neighbours' weights in row order in fp32, exactly as the oracle's
or_sweep_csr does, and draws the same Philox thresholds, so final states,
best states and the running best energies must equal the oracle's bit for
bit on every instance of a batch, whatever its size mix.  The FP64 energies
of the returned states must match the energy function within 1e-9 relative.
"""
import numpy as np
import pytest

import onesolver_amd
import oracle
from onesolver_amd import instances

pytestmark = pytest.mark.gpu


def _tscale(S, bmin=0.1, bmax=3.0):
    return onesolver_amd.tscale_from_beta(np.geomspace(bmin, bmax, S))


def _irregular(n, seed):
    """Symmetric CSR with isolated spins, a hub of high degree and real weights."""
    rng = np.random.default_rng(seed)
    pairs = set()
    hub = n // 2
    for l in rng.choice(n, size=min(n - 1, 13), replace=False):
        if l != hub:
            pairs.add((min(hub, l), max(hub, l)))
    for _ in range(n):
        a, b = rng.integers(0, n, 2)
        if a != b and a % 7 and b % 7:  # spins 0, 7, 14, ... stay isolated except the hub
            pairs.add((min(a, b), max(a, b)))
    rows, cols, vals = [], [], []
    for (a, b) in sorted(pairs):
        w = float(np.float32(rng.normal()))
        rows += [a, b]
        cols += [b, a]
        vals += [w, w]
    rows, cols, vals = np.array(rows), np.array(cols), np.array(vals)
    order = np.lexsort((cols, rows))
    rows, cols, vals = rows[order], cols[order], vals[order]
    rp = np.zeros(n + 1, dtype=np.int32)
    np.add.at(rp, rows + 1, 1)
    rp = np.cumsum(rp).astype(np.int32)
    diag = (rng.random(n, dtype=np.float32) - np.float32(0.5)).astype(np.float64)
    return rp, cols.astype(np.int32), vals.astype(np.float64), diag


def _batch():
    insts = [instances.regular_csr(n, d, seed=5, instance=i)
             for i, (n, d) in enumerate([(6, 3), (32, 5), (33, 4), (64, 8), (100, 3), (258, 7)])]
    insts.append(_irregular(70, 1))
    insts.append(_irregular(31, 2))
    return insts


def _check_against_oracle(out, insts, ts, R, seed, base, init=None):
    for i, (rp, cl, vl, dg) in enumerate(insts):
        ini = None if init is None else init[i]
        fin, bs, be = oracle.sweep_csr(rp, cl, vl, dg, ts, R, seed=seed, instance=base + i,
                                       init=ini)
        assert (out["final_states"][i] == fin).all(), f"final states differ, instance {i}"
        assert (out["best_states"][i] == bs).all(), f"best states differ, instance {i}"
        assert (out["run_best_energies"][i] == be).all(), f"running best differs, instance {i}"
        e64 = oracle.csr_energy(rp, cl, vl, dg, bs)
        np.testing.assert_allclose(out["best_energies"][i], e64, rtol=1e-9, atol=1e-12)
        r = int(np.argmin(e64))
        assert out["inst_best_replica"][i] == r
        assert out["inst_best_energy"][i] == out["best_energies"][i][r]


@pytest.mark.parametrize("use_graph", [True, False])
def test_csr_batch_matches_oracle_bit_for_bit(ctx, use_graph):
    insts = _batch()
    R, S = 100, 25
    ts = _tscale(S)
    print(f"[gpu] csr batch of {len(insts)} R={R} S={S} graph={use_graph}", flush=True)
    b = onesolver_amd.CsrBatch(ctx, insts)
    out = b.anneal(R, ts, seed=11, instance_base=4, final_states=True, use_graph=use_graph)
    _check_against_oracle(out, insts, ts, R, 11, 4)
    w, g, _ = b.launch_shape()
    assert w in (1, 2, 4) and g >= 1
    b.close()


def _partly_isolated(n, m, seed):
    """A 3-regular graph on the first m spins, the other n - m spins isolated,
    real weights."""
    rp0, cl0, vl0, dg0 = instances.regular_csr(m, 3, seed=seed, instance=0)
    rng = np.random.default_rng(seed)
    vl = np.asarray(vl0, dtype=np.float64).copy()
    # symmetric real weights: one draw per unordered pair
    w = {}
    for j in range(m):
        for k in range(rp0[j], rp0[j + 1]):
            key = (min(j, cl0[k]), max(j, cl0[k]))
            if key not in w:
                w[key] = float(np.float32(rng.normal()))
            vl[k] = w[key]
    rp = np.concatenate([rp0, np.full(n - m, rp0[-1], dtype=np.int32)]).astype(np.int32)
    dg = np.concatenate([np.asarray(dg0, np.float64),
                         (rng.random(n - m, dtype=np.float32) - np.float32(0.5)).astype(np.float64)])
    return rp, np.asarray(cl0, np.int32), vl, dg


@pytest.mark.parametrize("tune", [False, True])
def test_csr_prefetch_kernel_matches_oracle(ctx, tune):
    """Batches whose rows all have at most 8 couplings run the row-prefetching
    kernel (the default when untuned): n up to 2048, n != 0 mod 32, one- and
    two-chunk rows, isolated spins, R != 0 mod 64."""
    insts = [instances.regular_csr(n, d, seed=7, instance=i)
             for i, (n, d) in enumerate([(6, 3), (33, 4), (64, 8), (2048, 5), (1999, 6)])]
    insts.append(_partly_isolated(45, 20, 3))
    R, S = 100, 6
    ts = _tscale(S)
    print(f"[gpu] csr prefetch-kernel batch, tune={tune}", flush=True)
    b = onesolver_amd.CsrBatch(ctx, insts)
    out = b.anneal(R, ts, seed=21, instance_base=2, final_states=True, tune=tune)
    _check_against_oracle(out, insts, ts, R, 21, 2)
    w, g, pf = b.launch_shape()
    if not tune:
        assert pf == 1
    b.close()


def test_csr_graph_replays_with_new_parameters(ctx):
    insts = _batch()
    R, S = 64, 12
    b = onesolver_amd.CsrBatch(ctx, insts)
    ts = _tscale(S)
    first = b.anneal(R, ts, seed=3, final_states=True)
    ts2 = _tscale(S, 0.5, 5.0)
    print("[gpu] csr graph replay, new seed and schedule", flush=True)
    second = b.anneal(R, ts2, seed=4, final_states=True)
    _check_against_oracle(second, insts, ts2, R, 4, 0)
    third = b.anneal(R, ts, seed=3, final_states=True)
    for k in ("best_energies", "run_best_energies"):
        assert (first[k] == third[k]).all()
    for a, c in zip(first["final_states"], third["final_states"]):
        assert (a == c).all()
    b.close()


def test_csr_start_states(ctx):
    insts = _batch()
    R, S = 70, 9
    ts = _tscale(S)
    rng = np.random.default_rng(0)
    init = [rng.integers(0, 2, (R, len(d)), dtype=np.uint8) for (_, _, _, d) in insts]
    print("[gpu] csr start states", flush=True)
    b = onesolver_amd.CsrBatch(ctx, insts)
    out = b.anneal(R, ts, seed=2, init_states=init, final_states=True)
    _check_against_oracle(out, insts, ts, R, 2, 0, init=init)
    # zero sweeps: the best state is the start state
    z = b.anneal(R, ts[:0], seed=2, init_states=init, final_states=True)
    for i in range(len(insts)):
        assert (z["best_states"][i] == init[i]).all()
    b.close()


def test_csr_c3_batch_energies(ctx):
    """The 512-instance batch (n in [64, 2048], d in 3..8) at a few sweeps:
    every returned energy is the FP64 energy of its state, and one instance is
    checked against the oracle bit for bit."""
    insts = instances.batch_c3(512)
    R, S = 256, 4
    ts = _tscale(S)
    print(f"[gpu] csr C3 batch R={R} S={S}", flush=True)
    b = onesolver_amd.CsrBatch(ctx, insts)
    out = b.anneal(R, ts, seed=1)
    for i in range(0, 512, 37):
        rp, cl, vl, dg = insts[i]
        e64 = oracle.csr_energy(rp, cl, vl, dg, out["best_states"][i])
        np.testing.assert_allclose(out["best_energies"][i], e64, rtol=1e-9, atol=1e-12)
    i = int(np.argmax([len(d) for (_, _, _, d) in insts]))
    rp, cl, vl, dg = insts[i]
    fin, bs, be = oracle.sweep_csr(rp, cl, vl, dg, ts, R, seed=1, instance=i)
    assert (out["best_states"][i] == bs).all()
    assert (out["run_best_energies"][i] == be).all()
    b.close()


def test_csr_rejects_bad_input(ctx):
    rp = np.array([0, 1, 2], dtype=np.int32)
    with pytest.raises(onesolver_amd.OneSolverError):
        onesolver_amd.CsrBatch(ctx, [(rp, np.array([1, 5], dtype=np.int32),
                                      np.array([1.0, 1.0]), np.zeros(2))])
    with pytest.raises(onesolver_amd.OneSolverError):
        onesolver_amd.CsrBatch(ctx, [(np.array([1, 1, 2], dtype=np.int32),
                                      np.array([1, 0], dtype=np.int32), np.array([1.0, 1.0]),
                                      np.zeros(2))])


def _check_colour_against_oracle(b, out, insts, ts, R, seed, base):
    """site_order 1: the library's class order is the oracle's greedy
    colouring; final and best states equal or_sweep_csr with that visiting
    order bit for bit; running best energies agree to FP64 rounding (the
    kernel sums a sweep's dE per lane, then over lanes) and exactly where
    every dE is exact."""
    for i, (rp, cl, vl, dg) in enumerate(insts):
        order, cs = b.site_order(i)
        o2, cs2 = oracle.csr_colour_order(rp, cl)
        assert (order == o2).all() and (cs == cs2).all(), f"colouring differs, instance {i}"
        fin, bs, be = oracle.sweep_csr(rp, cl, vl, dg, ts, R, seed=seed, instance=base + i,
                                       order=order)
        assert (out["final_states"][i] == fin).all(), f"final states differ, instance {i}"
        assert (out["best_states"][i] == bs).all(), f"best states differ, instance {i}"
        np.testing.assert_allclose(out["run_best_energies"][i], be, rtol=1e-9, atol=1e-12)
        e64 = oracle.csr_energy(rp, cl, vl, dg, bs)
        np.testing.assert_allclose(out["best_energies"][i], e64, rtol=1e-9, atol=1e-12)


@pytest.mark.parametrize("sites", [4, 8, 16])
def test_csr_colour_classes_match_oracle(ctx, monkeypatch, sites):
    """The colour-class kernel with 4, 8 and 16 sites per step on a mixed
    batch: regular graphs, hubs, isolated spins, n != 0 mod 32, R != 0 mod 64."""
    insts = _batch() + [instances.regular_csr(2048, 5, seed=7, instance=9)]
    R, S = 100, 7
    ts = _tscale(S)
    monkeypatch.setenv("OS_CSR_COLOUR_SITES", str(sites))
    print(f"[gpu] csr colour classes, {sites} sites per step", flush=True)
    b = onesolver_amd.CsrBatch(ctx, insts)
    out = b.anneal(R, ts, seed=13, instance_base=3, final_states=True, site_order=1)
    _check_colour_against_oracle(b, out, insts, ts, R, 13, 3)
    assert b.colour_shape()[2] == sites
    b.close()


def test_csr_batch_create_reports_the_first_bad_instance(ctx):
    """Batch creation checks the instances on several host threads; the error
    is the first bad instance's in order (row_ptr checks before column checks
    within an instance), as a serial pass reports it, and good batches still
    build the greedy colouring the oracle gives."""
    good = [_irregular(40, k) for k in range(6)]

    def bad_col(inst):
        rp, cl, vl, dg = (np.array(a) for a in inst)
        cl[len(cl) // 2] = len(dg) + 3
        return rp, cl, vl, dg

    def bad_rp(inst):
        rp, cl, vl, dg = (np.array(a) for a in inst)
        rp[5], rp[6] = rp[6], rp[5] - 1
        return rp, cl, vl, dg

    cases = [
        (good[:2] + [bad_col(good[2])] + good[3:4] + [bad_rp(good[4])], "column index out of range"),
        (good[:1] + [bad_rp(good[1])] + [bad_col(good[2])], "row_ptr must be non-decreasing"),
        (good[:3] + [(np.zeros(1, np.int32), np.zeros(0, np.int32), np.zeros(0), np.zeros(0))],
         "instance size must be positive"),
    ]
    for insts, msg in cases:
        with pytest.raises(onesolver_amd.InvalidArgument, match=msg):
            onesolver_amd.CsrBatch(ctx, insts)
    b = onesolver_amd.CsrBatch(ctx, good)
    for i, (rp, cl, vl, dg) in enumerate(good):
        order, starts = b.site_order(i)
        o2, cs2 = oracle.csr_colour_order(rp, cl)
        assert (order == o2).all() and (starts == cs2).all()
    b.close()


def test_csr_colour_mixed_sites_per_step_match_oracle(ctx, monkeypatch):
    """The untuned policy with at most 128 steps per class sweep
    (OS_CSR_COLOUR_LIMIT) gives small instances 4 sites per step, mid-size 8
    and large 16, in one launch; every instance still follows the oracle's
    class-order chains."""
    monkeypatch.setenv("OS_CSR_COLOUR_LIMIT", "128")
    insts = [instances.regular_csr(n, d, seed=3, instance=i)
             for i, (n, d) in enumerate([(64, 3), (700, 5), (1500, 8), (300, 4), (2048, 3)])]
    insts.append(_irregular(70, 1))
    R, S = 64, 5
    ts = _tscale(S)
    print("[gpu] csr colour classes, mixed sites per step", flush=True)
    b = onesolver_amd.CsrBatch(ctx, insts)
    out = b.anneal(R, ts, seed=17, instance_base=2, final_states=True, site_order=1, tune=False)
    assert list(b.colour_sites()) == [4, 8, 16, 4, 16, 4]
    assert b.colour_shape()[2] == 0
    _check_colour_against_oracle(b, out, insts, ts, R, 17, 2)
    b.close()


def test_csr_colour_classes_exact_energies_and_tuning(ctx):
    """Integer couplings and quarter-step diagonals: the running energies are
    exact, so they equal the oracle's; the tuner picks a shape, and the
    sequential order on the same batch still gives the sequential chains."""
    insts = []
    for i, (n, d) in enumerate([(64, 3), (333, 6), (1024, 8), (47, 4)]):
        rp, cl, vl, dg = instances.regular_csr(n, d, seed=2, instance=i)
        insts.append((rp, cl, vl, np.round(np.asarray(dg, np.float64) * 4) / 4))
    R, S = 64, 9
    ts = _tscale(S)
    print("[gpu] csr colour classes, exact energies, tuned", flush=True)
    b = onesolver_amd.CsrBatch(ctx, insts)
    out = b.anneal(R, ts, seed=5, final_states=True, site_order=1, tune=True)
    for i, (rp, cl, vl, dg) in enumerate(insts):
        order, _ = b.site_order(i)
        fin, bs, be = oracle.sweep_csr(rp, cl, vl, dg, ts, R, seed=5, instance=i, order=order)
        assert (out["final_states"][i] == fin).all() and (out["best_states"][i] == bs).all()
        assert (out["run_best_energies"][i] == be).all()
    w, g, sites = b.colour_shape()
    assert w in (1, 2, 4) and g >= 1 and sites in (0, 4, 8, 16)
    assert set(b.colour_sites()) <= {4, 8, 16}
    seq = b.anneal(R, ts, seed=5, final_states=True)
    _check_against_oracle(seq, insts, ts, R, 5, 0)
    b.close()


# --------------------------------------------------------------------------
# population annealing on the persistent sparse kernels
# --------------------------------------------------------------------------


def _exact_batch():
    """Integer couplings and quarter-step diagonals: every dE and energy is
    exact, so running energies do not depend on summation order."""
    insts = []
    for i, (n, d) in enumerate([(64, 3), (333, 6), (1024, 8), (47, 4)]):
        rp, cl, vl, dg = instances.regular_csr(n, d, seed=4, instance=i)
        insts.append((rp, cl, vl, np.round(np.asarray(dg, np.float64) * 4) / 4))
    return insts


def _check_population(out, insts, ts, theta, R, seed, base, orders=None):
    for i, (rp, cl, vl, dg) in enumerate(insts):
        od = None if orders is None else orders[i]
        fin, bs, be, anc = oracle.population_csr(rp, cl, vl, dg, ts, theta, R, seed=seed,
                                                 instance=base + i, order=od, threads=8)
        assert (out["ancestors"][i] == anc).all(), f"ancestors differ, instance {i}"
        assert (out["final_states"][i] == fin).all(), f"final states differ, instance {i}"
        assert (out["best_states"][i] == bs).all(), f"best states differ, instance {i}"
        e64 = oracle.csr_energy(rp, cl, vl, dg, bs)
        np.testing.assert_allclose(out["best_energies"][i], e64, rtol=1e-9, atol=1e-12)
        assert min(len(np.unique(a)) for a in anc) < R  # resampling selected


@pytest.mark.parametrize("use_graph", [True, False])
def test_csr_population_matches_oracle(ctx, use_graph):
    """Sequential site order on the mixed batch (real weights, hubs, isolated
    spins): running energies are the oracle's sums, so ancestors, final and
    best states equal or_population_csr bit for bit."""
    insts = _batch()
    R, theta = 100, 3
    ts = _tscale(5, 0.2, 3.0)
    print(f"[gpu] csr population, sequential order, graph={use_graph}", flush=True)
    b = onesolver_amd.CsrBatch(ctx, insts)
    out = b.population(R, ts, theta, seed=21, instance_base=4, final_states=True,
                       use_graph=use_graph)
    _check_population(out, insts, ts, theta, R, 21, 4)
    b.close()


def test_csr_population_colour_classes_match_oracle(ctx):
    insts = _exact_batch()
    R, theta = 64, 2
    ts = _tscale(6, 0.3, 3.0)
    print("[gpu] csr population, colour classes", flush=True)
    b = onesolver_amd.CsrBatch(ctx, insts)
    out = b.population(R, ts, theta, seed=8, final_states=True, site_order=1)
    orders = [b.site_order(i)[0] for i in range(len(insts))]
    _check_population(out, insts, ts, theta, R, 8, 0, orders)
    b.close()


def test_csr_population_at_equal_temperatures_is_plain_annealing(ctx):
    insts = _exact_batch()
    R, K, theta = 64, 4, 3
    print("[gpu] csr population at equal temperatures", flush=True)
    b = onesolver_amd.CsrBatch(ctx, insts)
    a = b.population(R, np.full(K, 0.7, np.float32), theta, seed=2, final_states=True)
    p = b.anneal(R, np.full(K * theta, 0.7, np.float32), seed=2, final_states=True)
    b.close()
    assert (a["ancestors"] == np.arange(R)).all()
    for i in range(len(insts)):
        assert (a["final_states"][i] == p["final_states"][i]).all()
        assert (a["best_states"][i] == p["best_states"][i]).all()
    assert (a["best_energies"] == p["best_energies"]).all()


def test_csr_population_rejects_bad_schedules(ctx):
    b = onesolver_amd.CsrBatch(ctx, _exact_batch()[:1])
    ts = np.array([2.0, 1.0], np.float32)
    with pytest.raises(onesolver_amd.OneSolverError):
        b.population(64, ts, 0)
    with pytest.raises(onesolver_amd.OneSolverError):
        b.population(64, np.zeros(0, np.float32), 2)
    with pytest.raises(onesolver_amd.OneSolverError):
        b.population(64, np.array([2.0, 0.0], np.float32), 2)
    with pytest.raises(onesolver_amd.OneSolverError):
        b.population(64, ts, 2, site_order=2)
    b.close()


def test_closing_a_context_closes_what_was_created_on_it():
    """Objects created on a context use its stream until destroyed; closing
    the context destroys them first, so none outlives the os_ctx."""
    c = onesolver_amd.Context(0)
    b = onesolver_amd.CsrBatch(c, _exact_batch()[:1])
    p = onesolver_amd.DenseProblem(c, np.eye(8))
    print("[gpu] context close with live children", flush=True)
    c.close()
    assert b._h is None and p._h is None
    b.close()
    p.close()
    with pytest.raises(onesolver_amd.OneSolverError):
        onesolver_amd.CsrBatch(c, _exact_batch()[:1])


# --------------------------------------------------------------------------
# replica exchange on the persistent sparse kernels
# --------------------------------------------------------------------------


@pytest.mark.parametrize("use_graph", [True, False])
def test_csr_tempering_matches_oracle(ctx, use_graph):
    """Sequential site order on the mixed batch: running energies are the
    oracle's sums, so final slots, final and best states equal
    or_tempering_csr bit for bit (a partial last segment included)."""
    insts = _batch()
    K, L, S, iv = 4, 16, 11, 3
    ts = _tscale(K, 0.3, 3.0)
    print(f"[gpu] csr tempering, sequential order, graph={use_graph}", flush=True)
    b = onesolver_amd.CsrBatch(ctx, insts)
    out = b.tempering(L, ts, S, iv, seed=23, instance_base=6, final_states=True,
                      use_graph=use_graph)
    b.close()
    for i, (rp, cl, vl, dg) in enumerate(insts):
        fin, bs, be, sl = oracle.tempering_csr(rp, cl, vl, dg, ts, L, S, iv, seed=23,
                                               instance=6 + i, threads=8)
        assert (out["final_slots"][i] == sl).all(), f"slots differ, instance {i}"
        assert (out["final_states"][i] == fin).all(), f"final states differ, instance {i}"
        assert (out["best_states"][i] == bs).all(), f"best states differ, instance {i}"
        np.testing.assert_allclose(out["best_energies"][i], oracle.csr_energy(rp, cl, vl, dg, bs),
                                   rtol=1e-9, atol=1e-12)
    moved = sum(int((out["final_slots"][i] != np.arange(K * L) % K).sum())
                for i in range(len(insts)))
    assert moved > 0  # exchanges happened


def test_csr_tempering_colour_classes_match_oracle(ctx):
    insts = _exact_batch()
    K, L, S, iv = 8, 8, 8, 2
    ts = _tscale(K, 0.3, 3.0)
    print("[gpu] csr tempering, colour classes", flush=True)
    b = onesolver_amd.CsrBatch(ctx, insts)
    out = b.tempering(L, ts, S, iv, seed=4, final_states=True, site_order=1)
    for i, (rp, cl, vl, dg) in enumerate(insts):
        order = b.site_order(i)[0]
        fin, bs, be, sl = oracle.tempering_csr(rp, cl, vl, dg, ts, L, S, iv, seed=4, instance=i,
                                               order=order, threads=8)
        assert (out["final_slots"][i] == sl).all(), f"slots differ, instance {i}"
        assert (out["final_states"][i] == fin).all(), f"final states differ, instance {i}"
        assert (out["best_states"][i] == bs).all(), f"best states differ, instance {i}"
    b.close()


def test_csr_tempering_with_equal_temperatures_is_plain_annealing(ctx):
    insts = _exact_batch()
    K, L, S = 4, 16, 9
    print("[gpu] csr tempering at equal temperatures", flush=True)
    b = onesolver_amd.CsrBatch(ctx, insts)
    a = b.tempering(L, np.full(K, 0.9, np.float32), S, 2, seed=7, final_states=True)
    p = b.anneal(K * L, np.full(S, 0.9, np.float32), seed=7, final_states=True)
    b.close()
    for i in range(len(insts)):
        assert (a["final_states"][i] == p["final_states"][i]).all()
        assert (a["best_states"][i] == p["best_states"][i]).all()
    assert (a["best_energies"] == p["best_energies"]).all()


def test_csr_tempering_rejects_bad_shapes(ctx):
    b = onesolver_amd.CsrBatch(ctx, _exact_batch()[:1])
    ts = np.array([2.0, 1.0], np.float32)
    with pytest.raises(onesolver_amd.OneSolverError):
        b.tempering(4, ts, 8, 0)                                   # no exchange interval
    with pytest.raises(onesolver_amd.OneSolverError):
        b.tempering(4, np.array([2.0], np.float32), 8, 2)          # one slot
    with pytest.raises(onesolver_amd.OneSolverError):
        b.tempering(4, np.array([2.0, -1.0], np.float32), 8, 2)
    with pytest.raises(onesolver_amd.OneSolverError):
        b.tempering(4, ts, 0, 2)                                   # no sweeps
    b.close()
