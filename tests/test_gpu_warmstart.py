"""Ablation of the learned warm start (GPU): what the trained network gains
over the method without it.

The network is trained on the GPU on SK instances (n=1024, the dense
benchmark distribution) labelled by the plain annealer, then compared on
unseen instances with
  * the uninformed predictor p = 1/2 and the untrained network (held-out
    cross-entropy: information about good states), and
  * plain random-init annealing on its own schedule (beta 0.1 -> 3.0) and
    random starts on the warm start's colder schedule (beta 1.0 -> 3.0), by
    time-to-99% of the best known energy.
The first comparison with plain annealing mixes two effects, the start
states and the colder schedule they allow; the comparison on the same
schedule isolates the network.  Every energy compared is the FP64 energy of
a returned state.
"""
import math

import numpy as np
import pytest

import onesolver_amd
from onesolver_amd import warmstart as ws

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def trained(ctx):
    net = onesolver_amd.WarmStartNet(ctx, seed=1)
    print("[gpu] warm start: training on 2048 SK instances n=1024", flush=True)
    losses, _ = ws.train(ctx, net, n=1024, instances=2048, epochs=2, lr=2e-3, seed=1)
    yield net, losses
    net.close()


def test_training_lowers_the_loss(trained):
    _, losses = trained
    assert np.mean(losses[-32:]) < np.mean(losses[:8]) - 0.02


def test_trained_network_predicts_unseen_instances_better_than_without_it(ctx, trained):
    net, _ = trained
    fresh = onesolver_amd.WarmStartNet(ctx, seed=1)
    print("[gpu] warm start: held-out cross-entropy", flush=True)
    l_trained = ws.heldout_loss(ctx, net, instances=64)
    l_fresh = ws.heldout_loss(ctx, fresh, instances=64)
    fresh.close()
    print(f"[gpu] held-out BCE trained {l_trained:.4f} untrained {l_fresh:.4f} "
          f"uninformed {math.log(2):.4f}", flush=True)
    assert l_trained < l_fresh
    assert l_trained < math.log(2) - 0.003


def test_colder_schedule_with_learned_start_reaches_target_no_later_than_plain(ctx, trained):
    """Learned start + colder schedule against plain annealing: this measures
    the schedule effect as much as the network (random starts on the colder
    schedule are reported beside it)."""
    net, _ = trained
    print("[gpu] warm start: time-to-99% ladder", flush=True)
    res = ws.evaluate(ctx, net, n=1024, instances=32, max_sweeps=1024, log=print)
    plain = res["plain"][1]
    learned = res["learned"][1]
    same = res["random_warm_schedule"][1]
    assert plain is not None and learned is not None and same is not None
    assert learned["sweeps"] <= plain["sweeps"]
    assert learned["wall_s"] <= plain["wall_s"] * 1.05
    # on the same schedule the learned start is reported beside random starts;
    # it must not need a later rung of the sweep ladder than twice theirs
    print(f"[gpu] same schedule: learned {learned['sweeps']} sweeps, random {same['sweeps']}",
          flush=True)
    assert learned["sweeps"] <= 2 * same["sweeps"]


# ---- mixed-sparse instances ------------------------------------------------


def test_csr_init_marginals_follow_the_dense_warm_start_rule(ctx):
    """os_csr_params::init_marginals samples the start planes on the GPU with
    the dense path's rule: the run equals the one started from the oracle's
    warm_init states, bit for bit."""
    import oracle
    from onesolver_amd import instances

    insts = [instances.regular_csr(n, d, seed=3, instance=i)
             for i, (n, d) in enumerate([(40, 3), (97 + 1, 4), (130, 5), (33 + 1, 6)])]
    rng = np.random.default_rng(4)
    marg = [rng.random(len(dg)).astype(np.float32) for (_, _, _, dg) in insts]
    R, seed, base = 64, 21, 5
    ts = onesolver_amd.tscale_from_beta(np.geomspace(0.5, 3.0, 3))
    init = [oracle.warm_init(p, R, seed=seed, instance=base + i) for i, p in enumerate(marg)]
    print("[gpu] csr init_marginals vs oracle warm_init", flush=True)
    b = onesolver_amd.CsrBatch(ctx, insts)
    for order in (0, 1):
        a = b.anneal(R, ts, seed=seed, instance_base=base, init_marginals=marg, final_states=True,
                     site_order=order)
        c = b.anneal(R, ts, seed=seed, instance_base=base, init_states=init, final_states=True,
                     site_order=order)
        for i in range(len(insts)):
            assert (a["final_states"][i] == c["final_states"][i]).all(), (order, i)
            assert (a["best_states"][i] == c["best_states"][i]).all(), (order, i)
        assert (a["best_energies"] == c["best_energies"]).all()
    b.close()


def test_sparse_warm_start_gains_at_a_fixed_budget(ctx):
    """The network on mixed-sparse (config-3 distribution) instances: trained
    on 256 instances labelled by the sparse annealer, its marginals seed
    unseen instances.  On the same schedule and sampler, starts from its
    marginals put more instances at 99% of the best known energy after a few
    cold sweeps than starts from p = 1/2."""
    from onesolver_amd import instances

    train = instances.batch_c3(256, seed=201, n_max=256)
    test = instances.batch_c3(64, seed=202, n_max=256)
    print("[gpu] sparse warm start: labels and training (256 instances)", flush=True)
    labels, _ = ws.sparse_labels(ctx, train, replicas=64, sweeps=512)
    net = onesolver_amd.WarmStartNet(ctx, seed=1)
    losses = ws.train_sparse(ctx, net, train, labels, epochs=3)
    assert np.mean(losses[-10:]) < np.mean(losses[:10])
    p = ws.sparse_marginals(ctx, net, test)
    net.close()
    b = onesolver_amd.CsrBatch(ctx, test)
    R = 128
    ref = b.anneal(R, ws.geometric(0.1, 3.0, 2048), seed=99, best_states=False, site_order=1)
    targets = 0.99 * ref["inst_best_energy"]
    half = [np.full(len(q), 0.5, dtype=np.float32) for q in p]
    ladder = [2, 4, 8]
    learned, _ = ws.sparse_time_to_target(b, targets, R, ladder, (3.0, 3.0), seed=11, marginals=p)
    blind, _ = ws.sparse_time_to_target(b, targets, R, ladder, (3.0, 3.0), seed=11,
                                        marginals=half)
    b.close()
    for x, y in zip(learned, blind):
        print(f"[gpu] {x['sweeps']} sweeps at beta 3: learned {x['frac_at_target']:.3f} "
              f"({x['sum_best']:.1f}), p=1/2 {y['frac_at_target']:.3f} ({y['sum_best']:.1f})",
              flush=True)
    assert learned[0]["sum_best"] < blind[0]["sum_best"]
    assert learned[1]["frac_at_target"] >= blind[1]["frac_at_target"] + 0.1


def test_sparse_inference_matches_the_densified_instances(ctx):
    """os_mpnn_predict_csr (sparse aggregation on the CSR batch) computes the
    network's marginals as on the densified instances it was trained on; the
    two differ only by the fp32 summation order of the aggregation."""
    from onesolver_amd import instances

    insts = instances.batch_c3(24, seed=203, n_max=600)
    net = onesolver_amd.WarmStartNet(ctx, seed=7)
    print("[gpu] sparse inference vs densified", flush=True)
    a = ws.sparse_marginals(ctx, net, insts)
    b = ws.sparse_marginals_dense(ctx, net, insts)
    net.close()
    diff = np.concatenate([np.abs(x - y) for x, y in zip(a, b)])
    print(f"[gpu] max |dp| {diff.max():.2e}, mean {diff.mean():.2e}", flush=True)
    assert diff.max() < 2e-2
    assert diff.mean() < 1e-3
