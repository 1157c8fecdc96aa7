"""Learned warm start: training and evaluation driver (GPU only).

Training data are SK instances of the dense benchmark distribution generated
on the GPU (os_dense_batch_create_sk) in chunks; their labels are the best
states of the plain GPU annealer.  Chunks are regenerated from their seeds
every epoch, so 20k N=1024 instances cost no host memory or transfer.

Evaluation follows the time-to-target definition: a method reaches the
target on an instance when its FP64-verified best energy is <= 0.99 E_ref
(E_ref < 0: the best energy of a long plain run); its time-to-99% on a batch
is the wall time of the shortest schedule (from a doubling ladder of sweep
counts) at which every instance of the batch reaches its target, including
the network's inference for the warm start.
"""
from __future__ import annotations

import time

import numpy as np

from . import Context, CsrBatch, DenseBatch, WarmStartNet, instances as _inst, tscale_from_beta


def geometric(beta0: float, beta1: float, sweeps: int) -> np.ndarray:
    return tscale_from_beta(np.geomspace(beta0, beta1, max(sweeps, 1)))[:sweeps]


def best_states(data: DenseBatch, replicas: int, sweeps: int, seed: int, instance: int,
                beta=(0.1, 3.0)):
    """Best state and FP64 energy of each instance after a plain anneal."""
    out = data.anneal(replicas, geometric(beta[0], beta[1], sweeps), seed=seed,
                      instance=instance, best_states=False)
    return out["best_state"], out["best_energy"]


def consensus_labels(data: DenseBatch, replicas: int, sweeps: int, seed: int, instance: int,
                     top: int = 16, beta=(0.1, 3.0)):
    """Majority vote of the `top` lowest-energy replicas of a plain anneal,
    each taken with spin 0 set to 1 (a state and its complement have the same
    SK energy): spins frozen across good states get a firm label."""
    out = data.anneal(replicas, geometric(beta[0], beta[1], sweeps), seed=seed,
                      instance=instance, best_states=True)
    st, e = out["best_states"], out["best_energies"]
    order = np.argsort(e, axis=1, kind="stable")[:, :top]
    sel = np.take_along_axis(st, order[:, :, None], axis=1)
    sel = np.where(sel[:, :, :1] == 1, sel, 1 - sel)
    return (sel.mean(axis=1) >= 0.5).astype(np.uint8), out["best_energy"]


def train(ctx: Context, net: WarmStartNet, n: int = 1024, instances: int = 20000,
          chunk: int = 256, epochs: int = 2, minibatch: int = 16, lr: float = 1e-3,
          seed: int = 1, base: int = 1 << 20, label_replicas: int = 64,
          label_sweeps: int = 512, label_top: int = 16, log=None):
    """Trains `net` on `instances` generated SK instances (ids base ...) and
    returns the per-step losses and the label-generation time."""
    chunks = [(c, min(chunk, instances - c)) for c in range(0, instances, chunk)]
    labels = {}
    t0 = time.perf_counter()
    for c, k in chunks:
        d = DenseBatch(ctx, sk=(k, n, seed, base + c))
        labels[c], _ = consensus_labels(d, label_replicas, label_sweeps, seed=seed + 17,
                                        instance=base + c, top=label_top)
        d.close()
    t_labels = time.perf_counter() - t0
    if log:
        log(f"labels for {instances} instances in {t_labels:.1f} s")
    losses = []
    for ep in range(epochs):
        for c, k in chunks:
            d = DenseBatch(ctx, sk=(k, n, seed, base + c))
            for m in range(0, k, minibatch):
                cnt = min(minibatch, k - m)
                losses.append(net.train_step(d, labels[c][m:m + cnt], lr, first=m))
            d.close()
        if log:
            log(f"epoch {ep}: mean loss of the last chunk {np.mean(losses[-max(1, k // minibatch):]):.4f}")
    return losses, t_labels


def heldout_loss(ctx: Context, net: WarmStartNet, n: int = 1024, instances: int = 64,
                 seed: int = 3, base: int = 1 << 23, label_replicas: int = 64,
                 label_sweeps: int = 512, label_top: int = 16):
    """Mean cross-entropy of the network's marginals against consensus labels
    of unseen instances (ln 2 = 0.693 for the uninformed p = 1/2)."""
    d = DenseBatch(ctx, sk=(instances, n, seed, base))
    y, _ = consensus_labels(d, label_replicas, label_sweeps, seed=seed + 17, instance=base,
                            top=label_top)
    loss, _ = net.loss_grad(d, y)
    d.close()
    return loss


def ladder(max_sweeps: int):
    s, out = 4, []
    while s <= max_sweeps:
        out.append(s)
        s *= 2
    return out


def time_to_target(data: DenseBatch, targets, replicas: int, sweeps_ladder, beta, seed,
                   instance, marginals=None, t_infer: float = 0.0):
    """Per rung: (sweeps, wall seconds incl. t_infer, fraction of instances at
    target, mean best energy); and the first rung where all reached it."""
    rows, hit = [], None
    for S in sweeps_ladder:
        ts = geometric(beta[0], beta[1], S)
        t0 = time.perf_counter()
        out = data.anneal(replicas, ts, seed=seed, instance=instance, best_states=False,
                          init_marginals=marginals)
        wall = time.perf_counter() - t0 + t_infer
        e = out["best_energy"]
        frac = float(np.mean(e <= targets))
        rows.append({"sweeps": S, "wall_s": wall, "frac_at_target": frac,
                     "mean_best": float(np.mean(e))})
        if hit is None and frac == 1.0:
            hit = rows[-1]
    return rows, hit


def evaluate(ctx: Context, net: WarmStartNet, n: int = 1024, instances: int = 32,
             replicas: int = 256, seed: int = 2, base: int = 1 << 22, ref_sweeps: int = 4096,
             max_sweeps: int = 1024, warm_beta=(1.0, 3.0), log=None):
    """Time-to-99% of plain SA (random init, beta 0.1 -> 3.0), of random init
    on the warm schedule, of the untrained network and of `net`."""
    data = DenseBatch(ctx, sk=(instances, n, seed, base))
    _, e_ref = best_states(data, replicas, ref_sweeps, seed=seed + 99, instance=base)
    targets = 0.99 * e_ref
    lad = ladder(max_sweeps)
    res = {"instances": instances, "n": n, "replicas": replicas, "ref_sweeps": ref_sweeps,
           "e_ref_mean": float(np.mean(e_ref))}
    # the call is repeated once untimed so every rung runs warm
    data.anneal(replicas, geometric(0.1, 3.0, 4), seed=seed, instance=base, best_states=False)
    res["plain"] = time_to_target(data, targets, replicas, lad, (0.1, 3.0), seed, base)
    res["random_warm_schedule"] = time_to_target(data, targets, replicas, lad, warm_beta, seed,
                                                 base)
    fresh = WarmStartNet(ctx, seed=12345)
    p0 = fresh.predict(data)
    res["untrained_net"] = time_to_target(data, targets, replicas, lad, warm_beta, seed, base,
                                          marginals=p0)
    fresh.close()
    net.predict(data)  # warm-up
    t0 = time.perf_counter()
    p = net.predict(data)
    t_inf = time.perf_counter() - t0
    res["t_infer_s"] = t_inf
    res["learned"] = time_to_target(data, targets, replicas, lad, warm_beta, seed, base,
                                    marginals=p, t_infer=t_inf)
    res["marginal_confidence"] = float(np.mean(np.abs(p - 0.5)))
    data.close()
    if log:
        for k in ("plain", "random_warm_schedule", "untrained_net", "learned"):
            rows, hit = res[k]
            rungs = " ".join(f"{r['sweeps']}:{r['frac_at_target']:.2f}" for r in rows)
            tail = (f"-> {hit['wall_s'] * 1e3:.1f} ms at {hit['sweeps']} sweeps" if hit
                    else "-> not reached")
            log(f"{k:22s} {rungs} {tail}")
    return res


# ---- mixed-sparse instances ------------------------------------------------
# The network reads dense instances.  A sparse instance (CSR) is densified
# with one extra spin in front: spin 0 is isolated and labelled 1, so the
# network's canonical labels (spin 0 pinned to 1) are the instance's own
# states, which is what instances with a diagonal field need (their
# complement is not a symmetry).  Instances are zero-padded to 128-multiples
# and batched by that size; padding spins are isolated and labelled 0.


def pinned_size(n: int) -> int:
    return (n + 1 + 127) // 128 * 128


def pinned_dense(inst, size: int) -> np.ndarray:
    """[size, size] FP64: spin j of the instance at j + 1, spin 0 isolated."""
    rp, cl, vl, dg = inst
    n = len(dg)
    q = np.zeros((size, size))
    q[1:n + 1, 1:n + 1] = _inst.csr_to_dense(rp, cl, vl, dg)
    return q


def _spins(item) -> int:
    """n of a CSR tuple (row_ptr, col, val, diag) or of a dense [n, n] matrix."""
    return item.shape[0] if isinstance(item, np.ndarray) else len(item[3])


def _pinned(item, size: int) -> np.ndarray:
    if isinstance(item, np.ndarray):
        q = np.zeros((size, size))
        n = item.shape[0]
        q[1:n + 1, 1:n + 1] = item
        return q
    return pinned_dense(item, size)


def _by_size(insts, chunk: int):
    groups = {}
    for i, inst in enumerate(insts):
        groups.setdefault(pinned_size(_spins(inst)), []).append(i)
    for size in sorted(groups):
        ids = groups[size]
        for c in range(0, len(ids), chunk):
            yield size, ids[c:c + chunk]


def sparse_labels(ctx: Context, insts, replicas: int = 64, sweeps: int = 512, seed: int = 7,
                  instance_base: int = 0, site_order: int = 1):
    """The best state (lowest FP64 energy over the replicas) of a plain
    anneal of every instance, and its energy."""
    b = CsrBatch(ctx, insts)
    out = b.anneal(replicas, geometric(0.1, 3.0, sweeps), seed=seed, instance_base=instance_base,
                   best_states=True, site_order=site_order)
    b.close()
    best = [bs[int(np.argmin(be))] for bs, be in zip(out["best_states"], out["best_energies"])]
    return best, out["inst_best_energy"]


def train_sparse(ctx: Context, net: WarmStartNet, insts, labels, epochs: int = 2,
                 chunk: int = 32, minibatch: int = 8, lr: float = 1e-3, log=None):
    """Trains `net` on instances with their label states, in the pinned dense
    layout; an instance is a CSR tuple (row_ptr, col, val, diag) or a dense
    [n, n] matrix with a diagonal field.  Returns the per-step losses."""
    losses = []
    for ep in range(epochs):
        for size, ids in _by_size(insts, chunk):
            d = DenseBatch(ctx, np.stack([_pinned(insts[i], size) for i in ids]))
            y = np.zeros((len(ids), size), dtype=np.uint8)
            y[:, 0] = 1
            for k, i in enumerate(ids):
                y[k, 1:_spins(insts[i]) + 1] = labels[i]
            for m in range(0, len(ids), minibatch):
                losses.append(net.train_step(d, y[m:m + minibatch], lr, first=m))
            d.close()
        if log:
            log(f"sparse epoch {ep}: mean loss {np.mean(losses):.4f}")
    return losses


def sparse_marginals(ctx: Context, net: WarmStartNet, insts, batch: CsrBatch = None):
    """Per-instance marginals [n_i] of the network, from the sparse batch
    (os_mpnn_predict_csr; `batch` holds `insts`, or one is made): the
    features and couplings the network saw in training, the scale over the
    padded dense size."""
    own = batch is None
    if own:
        batch = CsrBatch(ctx, insts)
    p = net.predict_csr(batch, [pinned_size(len(inst[3])) for inst in insts])
    if own:
        batch.close()
    return p


def sparse_marginals_dense(ctx: Context, net: WarmStartNet, insts, chunk: int = 32):
    """The same marginals through the pinned dense layout (the training
    layout): for checking the sparse inference, and for dense instances with
    a diagonal field (`insts` of [n, n] matrices)."""
    out = [None] * len(insts)
    for size, ids in _by_size(insts, chunk):
        d = DenseBatch(ctx, np.stack([_pinned(insts[i], size) for i in ids]))
        p = net.predict(d)
        d.close()
        for k, i in enumerate(ids):
            out[i] = np.ascontiguousarray(p[k, 1:_spins(insts[i]) + 1])
    return out


def sparse_time_to_target(batch: CsrBatch, targets, replicas: int, sweeps_ladder, beta, seed,
                          marginals=None, t_infer: float = 0.0, site_order: int = 1):
    """Per rung: (sweeps, wall seconds incl. t_infer, fraction of instances
    at target, sum of best energies); and the first rung where all reached
    it."""
    rows, hit = [], None
    for S in sweeps_ladder:
        t0 = time.perf_counter()
        out = batch.anneal(replicas, geometric(beta[0], beta[1], S), seed=seed,
                           init_marginals=marginals, best_states=False, site_order=site_order)
        wall = time.perf_counter() - t0 + t_infer
        e = out["inst_best_energy"]
        frac = float(np.mean(e <= targets))
        rows.append({"sweeps": S, "wall_s": wall, "frac_at_target": frac,
                     "sum_best": float(np.sum(e))})
        if hit is None and frac == 1.0:
            hit = rows[-1]
    return rows, hit
