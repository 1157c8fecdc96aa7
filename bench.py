"""Benchmark of the MI355X simulated-annealing QUBO solver.

Default workload (the configuration the headline metric is quoted on):
dense SK-type Ising mapped to QUBO, N=1024, J_ij ~ N(0,1)/sqrt(N), h=0,
16384 replicas, 4096 sweeps, geometric beta 0.1 -> 3.0 (inverse temperature),
bf16 hi+lo MFMA delayed field updates with block k=16, FP64 field refresh
every 512 sweeps.  One step = one full anneal of that batch with Q already
resident in HBM.  Prints one JSON line (see the keys below).

    python bench.py [--gpus 1] [--steps K] [--warmup W]
                    [--workload c1|c2|c3|c4|c5|exh|pt|ingest|batch|copy]

Other workloads: c1 dense N=28 in reference semantics (1000 iterations x
1024 tries); c3 a batch of 512 sparse instances; c4 the learned warm start
(time to 99% of the best known energy); c5 dense uniform N=8192 with the
fields in HBM; exh the exhaustive search (all 2^n states, dense U[-1,1] upper
triangle, n=32 by default); pt replica exchange on the config-2 instance
(256 ladders x 64 temperatures, plain annealing at the same budget beside it);
ingest the GPU `.qubo` reader on the text of a dense N=8192 instance already
in GPU memory (33.6M lines, bytes/s); batch one-solver-batch over the config-3
batch written as 512 `.qubo` files (instances/s, files to CSVs); copy the 1 GiB device copy that
calibrates the PMC traffic.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_BF16_TFLOPS = 2500.0  # MI355X dense bf16 MFMA
PEAK_HBM_GBS = 8000.0
PEAK_FP64_TFLOPS = 78.6    # MI355X FP64 vector

# HBM bytes per launch of the dominant kernel, from two rocprofv3 --pmc runs
# (FETCH_SIZE, WRITE_SIZE; KiB per dispatch) on the default configurations,
# calibrated on a 1 GiB device copy (read bytes = 2 x FETCH_SIZE, written
# bytes = WRITE_SIZE).  Only reported when the run uses that configuration.
TRAFFIC_DENSE_512 = (2 * 2967756.3 + 252689.6) * 1024   # dense_sweep_kernel, 512 sweeps
TRAFFIC_CSR_C3 = (2 * 9356018.9 + 7374460.7) * 1024     # csr_sweep_pf_kernel, 2048 sweeps
TRAFFIC_CSR_C3_COLOUR = (2 * 20880.6 + 4838077.6) * 1024  # csr_sweep_col_kernel, 2048 sweeps
# csr_sweep_col_kernel, one 2048-sweep launch of the config-3 batch (PMC passes): L2 requests and
# hits, LDS instructions (wave level), LDS bank-conflict and active cycles, VALU instructions
COUNTERS_CSR_C3_COLOUR = {"tcc_req": 6640911323.0, "tcc_hit": 6640351949.0,
                          "sq_insts_lds": 25849097501.0, "sq_lds_bank_conflict": 88772708309.0,
                          "sq_lds_idx_active": 143560519727.0, "sq_insts_valu": 274166665953.0}
TRAFFIC_PASS_C5 = (2 * 107418.5 + 131072.0) * 1024      # stream_pass_kernel, 5 blocks
TRAFFIC_INGEST_8192 = (2 * 643639.7 + 558661.2) * 1024   # parse_lines_kernel<true>, N=8192 text


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def _host_threads():
    threads = os.cpu_count() or 1
    try:
        threads = min(threads, len(os.sched_getaffinity(0)))
    except AttributeError:
        pass
    return min(threads, 16)  # the host share of one GPU


def cpu_baseline_dense(q, n, budget_s=8.0):
    """The reference's host paths restated: random-site attempts with an O(N^2)
    FP64 energy recompute each (or_anneal_literal) on one thread (the
    reference's --device-type host) and one replica per thread on all host
    cores (--device-type cpu), on bounded samples; plus the O(N) local-field
    restatements (reference semantics, and the sweep-mode port) on all cores."""
    import oracle

    beta = oracle.geometric_schedule(0.1, 3.0, 64)
    threads = _host_threads()
    # calibrate on a short run, then time about budget_s seconds of work
    t0 = time.perf_counter()
    oracle.anneal(q, n, beta[:8], 1, literal=True, threads=1)
    per = (time.perf_counter() - t0) / 8
    iters = int(max(16, min(200000, budget_s / max(per, 1e-9))))
    sched = np.full(iters, 1.0)
    t0 = time.perf_counter()
    oracle.anneal(q, n, sched, 1, literal=True, threads=1)
    lit_s = time.perf_counter() - t0
    lit_rate = iters / lit_s
    t0 = time.perf_counter()
    oracle.anneal(q, n, sched, threads, literal=True, threads=threads)
    litp_s = time.perf_counter() - t0
    # O(N) restatements: calibrated to about 2 s each
    t0 = time.perf_counter()
    oracle.anneal(q, n, np.full(2000, 1.0), threads, literal=False, threads=threads)
    per_f = (time.perf_counter() - t0) / 2000
    fld_iters = int(max(2000, min(20_000_000, 2.0 / max(per_f, 1e-12))))
    t0 = time.perf_counter()
    oracle.anneal(q, n, np.full(fld_iters, 1.0), threads, literal=False, threads=threads)
    fld_s = time.perf_counter() - t0
    reps = max(threads, 1) * 4
    t0 = time.perf_counter()
    oracle.sweep_dense(q, n, (1.0 / np.geomspace(0.1, 3.0, 1)).astype(np.float32), reps,
                       threads=threads)
    per_s = time.perf_counter() - t0
    nsw = int(max(1, min(256, 2.0 / max(per_s, 1e-9))))
    ts = (1.0 / np.geomspace(0.1, 3.0, nsw)).astype(np.float32)
    t0 = time.perf_counter()
    oracle.sweep_dense(q, n, ts, reps, threads=threads)
    f_s = time.perf_counter() - t0
    return {
        "value": lit_rate, "unit": "attempts/s", "cores": 1, "kind": "port",
        "sample": f"reference algorithm restated (O(N^2) FP64 energy per attempt, random site), "
                  f"1 replica x {iters} attempts on the N={n} instance, {lit_s:.1f} s",
        "all_cores": {"value": threads * iters / litp_s, "unit": "attempts/s", "cores": threads,
                      "sample": f"the same, {threads} replicas (one per thread) x {iters} "
                                f"attempts, {litp_s:.1f} s"},
        "reference_semantics_local_fields": {
            "value": threads * fld_iters / fld_s, "unit": "attempts/s", "cores": threads,
            "sample": f"reference semantics with O(N) FP64 local fields, {threads} replicas x "
                      f"{fld_iters} attempts, {fld_s:.1f} s"},
        "local_field_port": {"value": reps * len(ts) * n / f_s, "unit": "attempts/s",
                             "cores": threads,
                             "sample": f"sweep-mode O(N) local fields, {reps} replicas x "
                                       f"{len(ts)} sweeps, {f_s:.1f} s"},
    }


def run_dense(args):
    import onesolver_amd
    from onesolver_amd import instances

    n, R, S = args.n, args.replicas, args.sweeps
    log(f"generating SK instance n={n}")
    q32, const = instances.sk_qubo(n, seed=0, instance=0)
    q = q32.astype(np.float64)
    beta = np.geomspace(0.1, 3.0, S)
    ts = onesolver_amd.tscale_from_beta(beta)
    ctx = onesolver_amd.Context(0)
    log(f"device {ctx.device_name()}")
    prob = onesolver_amd.DenseProblem(ctx, q)
    kwargs = dict(seed=1234, refresh_interval=args.refresh, best_states=False)
    for i in range(args.warmup):
        out = prob.anneal(R, ts, **kwargs)
        log(f"warmup {i}: best {out['best_energy']:.4f} kernel {out['stats']['kernel_ms']:.1f} ms")
    kernel_ms, launches, drift = 0.0, 0, 0.0
    best = np.inf
    t0 = time.perf_counter()
    for i in range(args.steps):
        out = prob.anneal(R, ts, **kwargs)
        st = out["stats"]
        kernel_ms += st["kernel_ms"]
        launches += st["kernel_launches"]
        drift = max(drift, st["max_field_drift"])
        best = min(best, out["best_energy"])
        log(f"step {i}: best {out['best_energy']:.4f} (Ising {out['best_energy'] + const:.4f}) "
            f"kernel {st['kernel_ms']:.1f} ms total {st['total_ms']:.1f} ms")
    elapsed = time.perf_counter() - t0
    attempts = float(R) * S * n * args.steps
    value = attempts / elapsed
    sweeps_per_launch = S / (launches / args.steps)
    flops_per_launch = 4.0 * n * (R * sweeps_per_launch * n)
    avg_launch_s = kernel_ms / launches / 1e3
    achieved = flops_per_launch / avg_launch_s / 1e12
    res = {
        "metric": "spin-flip attempts/s", "value": value, "unit": "attempts/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "bf16 hi+lo MFMA, fp32 fields, fp64 energies",
        "data": "synthetic (Philox-generated SK instance, seed 0)",
        "config": {"workload": f"dense SK QUBO N={n}, {R} replicas, {S} sweeps, geometric beta "
                               f"0.1->3.0, block k=16, FP64 refresh every {args.refresh} sweeps",
                   "n": n, "replicas": R, "sweeps": S, "block_k": 16,
                   "refresh_interval": args.refresh, "parallelism": "replicas x columns on 1 GPU"},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": PEAK_BF16_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / PEAK_BF16_TFLOPS,
                     "traffic": TRAFFIC_DENSE_512 if (n, R, args.refresh) == (1024, 16384, 512)
                     else None,
                     "kernel": f"dense_sweep_kernel<{-(-n // 128)}>",
                     "avg_launch_ms": avg_launch_s * 1e3,
                     "algorithmic_flops_per_launch": flops_per_launch,
                     # the same work counted as FP32 (2N flop per attempt: one field update)
                     "fp32_equivalent_tflops": achieved / 2.0},
        "kernel_attempts_per_s": float(R) * S * n * args.steps / (kernel_ms / 1e3),
        "best_energy": best, "best_energy_ising": best + const,
        "max_field_drift": drift,
    }
    ctx.close()
    return res, q


def run_dense_c5(args):
    """Config 5: dense uniform QUBO N=8192 (Q fp32 U[-1,1), generated on the
    GPU), 4096 replicas, 1024 sweeps, blocks of k=64 spins with the fields in
    HBM (one MFMA pass over the fields per block), FP64 refresh every 256
    sweeps and FP64 energies of every returned state."""
    import onesolver_amd

    n, R, S = 8192, args.c5_replicas, args.c5_sweeps
    if args.c5_group > 0:
        os.environ["OS_STREAM_GROUP"] = str(args.c5_group)
    if args.c5_chunk > 0:
        os.environ["OS_STREAM_PASS_CHUNK"] = str(args.c5_chunk)
    if args.c5_wgs > 0:
        os.environ["OS_STREAM_PASS_WGS_PER_CU"] = str(args.c5_wgs)
    beta = np.geomspace(0.02, 2.0, S)
    ts = onesolver_amd.tscale_from_beta(beta)
    ctx = onesolver_amd.Context(0)
    log(f"device {ctx.device_name()}")
    b = onesolver_amd.DenseBatch(ctx, uniform=(1, n, 0, 0))
    kwargs = dict(seed=1234, refresh_interval=256, best_states=False)
    tuned = None
    if args.c5_group <= 0 and args.c5_wgs <= 0:
        # the autotuner picks blocks per pass and pass work-groups per CU on
        # probe runs of this batch (outside the timed steps)
        t_tune = time.perf_counter()
        tuned = b.tune(R, ts, seed=1234)
        tuned["tune_s"] = time.perf_counter() - t_tune
        log(f"tuned: {tuned}")
    for i in range(args.warmup):
        out = b.anneal(R, ts[:max(1, S // 8)], **kwargs)
        log(f"warmup {i}: kernel {out['stats']['kernel_ms']:.1f} ms")
    kernel_ms, launches, drift, best = 0.0, 0, 0.0, np.inf
    pass_ms, passes = 0.0, 0
    t0 = time.perf_counter()
    for i in range(args.steps):
        out = b.anneal(R, ts, **kwargs)
        st = out["stats"]
        kernel_ms += st["kernel_ms"]
        launches += st["kernel_launches"]
        pass_ms += st["pass_avg_ms"]
        passes += 1
        drift = max(drift, st["max_field_drift"])
        best = min(best, float(out["best_energy"][0]))
        log(f"step {i}: best {out['best_energy'][0]:.2f} kernel {st['kernel_ms']:.1f} ms "
            f"total {st['total_ms']:.1f} ms")
    elapsed = time.perf_counter() - t0
    attempts = float(R) * S * n * args.steps
    per_block_s = kernel_ms / launches / 1e3
    pass_s = pass_ms / max(passes, 1) / 1e3
    group = n // 64 * S / (out["stats"]["pass_launches"] or 1)
    field_bytes = 8.0 * R * n  # per pass: fields read and written once
    flops = 4.0 * n * R * 64 * group  # per pass: hi + lo products of `group` blocks
    b.close()
    ctx.close()
    return {
        "metric": "spin-flip attempts/s", "value": attempts / elapsed, "unit": "attempts/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "bf16 hi+lo MFMA, fp32 fields (HBM), fp64 energies",
        "data": "synthetic (Philox-generated uniform instance, seed 0)",
        "config": {"workload": f"dense uniform QUBO N={n}, {R} replicas, {S} sweeps, geometric "
                               f"beta 0.02->2.0, block k=64, fields in HBM, FP64 refresh every "
                               f"256 sweeps", "n": n, "replicas": R, "sweeps": S, "block_k": 64,
                   "tuned": tuned},
        "roofline": {"bound": "hbm", "achieved": field_bytes / pass_s / 1e9,
                     "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": field_bytes / pass_s / 1e9 / PEAK_HBM_GBS,
                     "traffic": TRAFFIC_PASS_C5 if R == 4096 and round(group) == 5 else None,
                     "kernel": "stream_pass_kernel", "avg_launch_ms": pass_s * 1e3,
                     "algorithmic_bytes_per_launch": field_bytes,
                     "blocks_per_pass": group, "mfma_tflops": flops / pass_s / 1e12},
        # the dense counting rule over the whole job: 4N bf16 flop per attempt
        "attempt_mfma": {"achieved": attempts / elapsed * 4.0 * n / 1e12,
                         "peak": PEAK_BF16_TFLOPS, "unit": "TFLOP/s",
                         "frac": attempts / elapsed * 4.0 * n / 1e12 / PEAK_BF16_TFLOPS},
        "ms_per_block": per_block_s * 1e3,
        "best_energy": best, "max_field_drift": drift,
    }, None


def run_reference_c1(args):
    """Config 1: N=28 dense, reference semantics, 1000 iterations x 1024 tries."""
    import onesolver_amd
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from qubo_io import load_qubo

    q = load_qubo(os.path.join(ROOT, "tests", "golden", "data", "c1_dense28.qubo"))
    beta = np.geomspace(0.1, 1.0, 1000)
    beta = np.array([0.1 * (10.0 ** (1.0 / 999)) ** i for i in range(1000)])
    ctx = onesolver_amd.Context(0)
    for _ in range(args.warmup):
        ctx.anneal_reference(q, beta, 1024, want_states=False)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        best, be, _, _ = ctx.anneal_reference(q, beta, 1024, want_states=False)
    elapsed = time.perf_counter() - t0
    attempts = 1024.0 * 1000 * args.steps
    ctx.close()
    return {
        "metric": "spin-flip attempts/s", "value": attempts / elapsed, "unit": "attempts/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp64", "data": "synthetic (committed N=28 instance)",
        "config": {"workload": "dense N=28, reference semantics, geometric 0.1->1.0, "
                               "1000 iterations x 1024 tries", "n": 28},
        "best_energy": be,
    }, q


def cpu_baseline_csr(inst, ts_full, budget_s=10.0):
    """The sparse sweep restated on the CPU (oracle.sweep_csr, the same
    fp32 gather per site), all host cores over replicas, on a bounded sample:
    one instance of the batch, a few sweeps."""
    import oracle

    rp, cl, vl, dg = inst
    threads = _host_threads()
    reps = max(threads, 1) * 8
    t0 = time.perf_counter()
    oracle.sweep_csr(rp, cl, vl, dg, ts_full[:2], reps, threads=threads)
    per = (time.perf_counter() - t0) / 2
    S = int(max(2, min(len(ts_full), budget_s / max(per, 1e-9))))
    t0 = time.perf_counter()
    oracle.sweep_csr(rp, cl, vl, dg, ts_full[:S], reps, threads=threads)
    el = time.perf_counter() - t0
    n = len(dg)
    return {"value": reps * S * n / el, "unit": "attempts/s", "cores": threads, "kind": "port",
            "sample": f"sparse sweep restated (fp32 gather per site), instance n={n} "
                      f"d={int(np.diff(rp).max())}, {reps} replicas x {S} sweeps, {el:.1f} s"}


def run_csr_c3(args):
    """Config 3: 512 sparse d-regular instances, n log-uniform in [64, 2048],
    256 replicas each, 2048 sweeps, one persistent kernel replayed from a
    hipGraph with the tuned launch shape.  --c3-order colour (default) visits
    each instance's sites class by class of a greedy colouring, several
    independent sites per wave step; sequential visits them 0..n-1."""
    import onesolver_amd
    from onesolver_amd import instances

    insts = instances.batch_c3(512)
    R, S = 256, args.sweeps_c3
    beta = np.geomspace(0.1, 3.0, S)
    ts = onesolver_amd.tscale_from_beta(beta)
    ctx = onesolver_amd.Context(0)
    log(f"device {ctx.device_name()}")
    b = onesolver_amd.CsrBatch(ctx, insts)
    order = 1 if args.c3_order == "colour" else 0
    shape = b.colour_shape if order else b.launch_shape
    for i in range(args.warmup):
        out = b.anneal(R, ts, seed=1234, best_states=False, tune=not args.c3_no_tune,
                       site_order=order)
        log(f"warmup {i}: graph {out['stats']['kernel_ms']:.1f} ms shape {shape()}")
    kernel_ms = 0.0
    t0 = time.perf_counter()
    for i in range(args.steps):
        out = b.anneal(R, ts, seed=1234 + i, best_states=False, tune=not args.c3_no_tune,
                       site_order=order)
        kernel_ms += out["stats"]["kernel_ms"]
        log(f"step {i}: graph {out['stats']['kernel_ms']:.1f} ms")
    elapsed = time.perf_counter() - t0
    ns = np.array([len(d) for (_, _, _, d) in insts], dtype=np.float64)
    deg = np.array([np.diff(rp).mean() for (rp, _, _, _) in insts])
    attempts = float(ns.sum()) * R * S * args.steps
    bytes_alg = float((ns * (8.0 * deg + 8.0)).sum()) * R * S  # per launch
    avg_launch_s = kernel_ms / args.steps / 1e3
    achieved = bytes_alg / avg_launch_s / 1e9
    res = {
        "metric": "spin-flip attempts/s", "value": attempts / elapsed, "unit": "attempts/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp32 fields, fp64 energies",
        "data": "synthetic (seeded random d-regular graphs)",
        "config": {"workload": f"batched sparse: 512 instances n in [64,2048] log-uniform, "
                               f"d-regular d in 3..8, weights +-1, {R} replicas, {S} sweeps, "
                               f"persistent kernel + hipGraph, site order {args.c3_order}",
                   "instances": 512, "replicas": R, "sweeps": S, "sum_n": int(ns.sum()),
                   "site_order": args.c3_order, "launch_shape": list(shape()),
                   "launch_shape_fields": ["waves per group", "groups",
                                           "sites per step" if order else "row prefetch"]},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": achieved / PEAK_HBM_GBS,
                     "traffic": (TRAFFIC_CSR_C3_COLOUR if order else TRAFFIC_CSR_C3)
                     if S == 2048 else None,
                     "kernel": "csr_sweep_col_kernel" if order else "csr_sweep_pf_kernel",
                     "avg_launch_ms": avg_launch_s * 1e3,
                     "algorithmic_bytes_per_launch": bytes_alg},
    }
    if order and S == 2048:
        # beside the algorithmic figure: what the counters of the same launch
        # give at this run's launch time
        k = COUNTERS_CSR_C3_COLOUR
        res["roofline"]["measured"] = {
            "hbm_GBs": TRAFFIC_CSR_C3_COLOUR / avg_launch_s / 1e9,
            "l2_requests_per_s": k["tcc_req"] / avg_launch_s,
            "l2_hit_rate": k["tcc_hit"] / k["tcc_req"],
            "lds_GBs": k["sq_insts_lds"] * 64 * 4 / avg_launch_s / 1e9,
            "lds_bank_conflict_cycle_frac": k["sq_lds_bank_conflict"] / k["sq_lds_idx_active"],
            "valu_lane_instructions_per_attempt": k["sq_insts_valu"] * 64 / (attempts / args.steps),
            "valu_wave_instructions_per_simd_cycle":
                k["sq_insts_valu"] / avg_launch_s / (256 * 4 * 2.4e9),
            "bound": "LDS latency per wave step (neighbour-bit gathers, 62% bank-conflict "
                     "cycles), see profiles/csr_sweep_col_kernel_sparse512.md"}
    if order:
        sites = b.colour_sites()
        res["config"]["instances_by_sites_per_step"] = {
            str(g): int((sites == g).sum()) for g in (4, 8, 16)}
    imax = int(np.argmax(ns))
    b.close()
    ctx.close()
    return res, (insts[imax], ts)


def run_warmstart_c4(args):
    """Config 4: learned warm start on the dense SK distribution (N=1024):
    train the network on the GPU on instances labelled by the plain annealer,
    then time-to-99% of the best known FP64 energy on unseen instances, warm
    start (inference included, 256 replicas) vs plain GPU annealing."""
    import onesolver_amd
    from onesolver_amd import warmstart as ws

    ctx = onesolver_amd.Context(0)
    log(f"device {ctx.device_name()}")
    net = onesolver_amd.WarmStartNet(ctx, seed=1)
    t0 = time.perf_counter()
    losses, t_labels = ws.train(ctx, net, n=1024, instances=args.c4_train, epochs=args.c4_epochs,
                                lr=2e-3, seed=1, log=log)
    t_train = time.perf_counter() - t0
    heldout = ws.heldout_loss(ctx, net, instances=64)
    log(f"trained in {t_train:.1f} s (labels {t_labels:.1f} s), held-out BCE {heldout:.4f}")
    res = ws.evaluate(ctx, net, n=1024, instances=args.c4_eval, max_sweeps=1024, log=log)
    plain, learned = res["plain"][1], res["learned"][1]
    ctx.close()
    out = {
        "metric": "time-to-99%-best energy (learned warm start)",
        "value": learned["wall_s"] if learned else None, "unit": "s",
        "n_gpus": 1, "steps": 1, "warmup": 1, "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None, "dtype": "bf16 weights fp32 acc (network); bf16 hi+lo MFMA (SA)",
        "data": "synthetic (SK instances generated on the GPU)",
        "config": {"workload": f"dense SK N=1024, network trained on {args.c4_train} instances "
                               f"x {args.c4_epochs} epochs, {args.c4_eval} unseen instances, "
                               f"256 replicas, sweep ladder 4..1024", "n": 1024},
        "plain_time_to_99_s": plain["wall_s"] if plain else None,
        "speedup_vs_plain": (plain["wall_s"] / learned["wall_s"]) if plain and learned else None,
        "random_init_warm_schedule_s": (res["random_warm_schedule"][1] or {}).get("wall_s"),
        "speedup_vs_random_same_schedule": (
            res["random_warm_schedule"][1]["wall_s"] / learned["wall_s"]
            if learned and res["random_warm_schedule"][1] else None),
        "untrained_net_s": (res["untrained_net"][1] or {}).get("wall_s"),
        "heldout_bce": heldout, "uninformed_bce": float(np.log(2.0)),
        "train_s": t_train, "label_s": t_labels, "infer_ms": res["t_infer_s"] * 1e3,
        "ladder": {k: res[k][0] for k in ("plain", "random_warm_schedule", "untrained_net",
                                          "learned")},
    }
    return out


def run_warmstart_c4_sparse(args):
    """The learned warm start on mixed-sparse instances: the network trained
    on config-3-distribution instances (densified with a pinned spin 0,
    labelled by the sparse annealer), then time-to-99% of the best known FP64
    energy on the config-3 batch (512 instances, 256 replicas, colour-class
    order): plain annealing, random and p = 1/2 starts on a cold schedule
    (beta 3), and the network's starts on that schedule (inference
    included)."""
    import onesolver_amd
    from onesolver_amd import instances
    from onesolver_amd import warmstart as ws

    ctx = onesolver_amd.Context(0)
    log(f"device {ctx.device_name()}")
    t0 = time.perf_counter()
    train = instances.batch_c3(args.c4_train, seed=101)
    labels, _ = ws.sparse_labels(ctx, train, replicas=64, sweeps=512)
    t_labels = time.perf_counter() - t0
    net = onesolver_amd.WarmStartNet(ctx, seed=1)
    t0 = time.perf_counter()
    losses = ws.train_sparse(ctx, net, train, labels, epochs=args.c4_epochs, log=log)
    t_train = time.perf_counter() - t0
    log(f"trained on {len(train)} instances in {t_train:.1f} s (labels {t_labels:.1f} s)")
    ev = instances.batch_c3(512, seed=0)
    b = onesolver_amd.CsrBatch(ctx, ev)
    R = 256
    ref = b.anneal(R, ws.geometric(0.1, 3.0, 4096), seed=99, best_states=False, site_order=1)
    targets = 0.99 * ref["inst_best_energy"]
    ws.sparse_marginals(ctx, net, ev, batch=b)  # warm-up
    t0 = time.perf_counter()
    p = ws.sparse_marginals(ctx, net, ev, batch=b)
    t_inf = time.perf_counter() - t0
    half = [np.full(len(q), 0.5, dtype=np.float32) for q in p]
    lad = ws.ladder(512)
    lad = [2] + lad
    b.anneal(R, ws.geometric(0.1, 3.0, 4), seed=11, best_states=False, site_order=1)
    res = {
        "plain": ws.sparse_time_to_target(b, targets, R, lad, (0.1, 3.0), seed=11),
        "random_cold": ws.sparse_time_to_target(b, targets, R, lad, (3.0, 3.0), seed=11),
        "half_cold": ws.sparse_time_to_target(b, targets, R, lad, (3.0, 3.0), seed=11,
                                              marginals=half),
        "learned_cold": ws.sparse_time_to_target(b, targets, R, lad, (3.0, 3.0), seed=11,
                                                 marginals=p, t_infer=t_inf),
    }
    b.close()
    net.close()
    ctx.close()
    for k, (rows, hit) in res.items():
        rungs = " ".join(f"{r['sweeps']}:{r['frac_at_target']:.3f}" for r in rows)
        log(f"{k:13s} {rungs} -> " + (f"{hit['wall_s'] * 1e3:.1f} ms at {hit['sweeps']} sweeps"
                                      if hit else "not reached"))
    hit = {k: v[1] for k, v in res.items()}
    at = {k: {r["sweeps"]: r["frac_at_target"] for r in v[0]} for k, v in res.items()}
    return {
        "metric": "time-to-99%-best energy (learned warm start, mixed-sparse)",
        "value": hit["learned_cold"]["wall_s"] if hit["learned_cold"] else None, "unit": "s",
        "n_gpus": 1, "steps": 1, "warmup": 1, "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None, "dtype": "bf16 weights fp32 acc (network); fp32 fields (SA)",
        "data": "synthetic (config-3 distribution, seeded d-regular graphs)",
        "config": {"workload": f"config-3 batch (512 instances, n 64..2048), network trained on "
                               f"{len(train)} instances x {args.c4_epochs} epochs, 256 replicas, "
                               f"colour-class order, sweep ladder 2..512", "path": "sparse"},
        "time_to_99_s": {k: (v["wall_s"] if v else None) for k, v in hit.items()},
        "sweeps_to_99": {k: (v["sweeps"] if v else None) for k, v in hit.items()},
        "frac_at_target_2_4_8_sweeps": {k: [at[k][2], at[k][4], at[k][8]] for k in at},
        "infer_s": t_inf, "train_s": t_train, "label_s": t_labels,
        "loss_first_last": [float(np.mean(losses[:10])), float(np.mean(losses[-10:]))],
    }


def run_warmstart_c5(args):
    """Config 5's learned vs plain init: the network trained on dense uniform
    instances (N=1024, Q U[-1,1), in the pinned dense layout since the
    diagonal breaks the complement symmetry; labels from the dense annealer),
    applied to the N=8192 instance of config 5; time-to-99% of the best known
    FP64 energy with 4096 replicas: plain init on the config-5 schedule
    (beta 0.02 -> 2.0), and random, p = 1/2 and learned starts on a colder
    schedule (beta 0.5 -> 2.0), inference included."""
    import onesolver_amd
    from onesolver_amd import warmstart as ws

    ctx = onesolver_amd.Context(0)
    log(f"device {ctx.device_name()}")
    t0 = time.perf_counter()
    qs, labels = [], []
    for c in range(0, args.c4_train, 64):
        k = min(64, args.c4_train - c)
        d = onesolver_amd.DenseBatch(ctx, uniform=(k, 1024, 5, 1000 + c))
        bs, _ = ws.best_states(d, 64, 256, seed=17, instance=1000 + c, beta=(0.02, 2.0))
        qs.extend(list(d.q()))
        labels.extend(list(bs))
        d.close()
    t_labels = time.perf_counter() - t0
    net = onesolver_amd.WarmStartNet(ctx, seed=1)
    t0 = time.perf_counter()
    losses = ws.train_sparse(ctx, net, qs, labels, epochs=args.c4_epochs, chunk=16, log=log)
    t_train = time.perf_counter() - t0
    del qs
    log(f"trained on {len(labels)} instances in {t_train:.1f} s (labels {t_labels:.1f} s)")
    n, R = 8192, args.c5_replicas
    b = onesolver_amd.DenseBatch(ctx, uniform=(1, n, 0, 0))
    q = b.q()[0]
    ws.sparse_marginals_dense(ctx, net, [q[:1024, :1024]])  # warm-up
    t0 = time.perf_counter()
    p = ws.sparse_marginals_dense(ctx, net, [q])[0]
    t_inf = time.perf_counter() - t0
    del q
    net.close()
    ref = b.anneal(R, ws.geometric(0.02, 2.0, 2048), seed=99, instance=0, best_states=False)
    targets = 0.99 * np.asarray(ref["best_energy"])
    lad = ws.ladder(1024)[2:]
    half = np.full((1, n), 0.5, dtype=np.float32)
    res = {
        "plain": ws.time_to_target(b, targets, R, lad, (0.02, 2.0), 11, 0),
        "random_warm": ws.time_to_target(b, targets, R, lad, (0.5, 2.0), 11, 0),
        "half_warm": ws.time_to_target(b, targets, R, lad, (0.5, 2.0), 11, 0, marginals=half),
        "learned_warm": ws.time_to_target(b, targets, R, lad, (0.5, 2.0), 11, 0,
                                          marginals=p[None, :], t_infer=t_inf),
    }
    b.close()
    ctx.close()
    for k, (rows, hit) in res.items():
        rungs = " ".join(f"{r['sweeps']}:{r['mean_best']:.0f}" for r in rows)
        log(f"{k:13s} {rungs} -> " + (f"{hit['wall_s']:.3f} s at {hit['sweeps']} sweeps"
                                      if hit else "not reached"))
    hit = {k: v[1] for k, v in res.items()}
    return {
        "metric": "time-to-99%-best energy (learned warm start, config 5)",
        "value": hit["learned_warm"]["wall_s"] if hit["learned_warm"] else None, "unit": "s",
        "n_gpus": 1, "steps": 1, "warmup": 1, "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None, "dtype": "bf16 weights fp32 acc (network); bf16 hi+lo MFMA (SA)",
        "data": "synthetic (uniform instances generated on the GPU)",
        "config": {"workload": f"dense uniform N={n}, {R} replicas; network trained on "
                               f"{len(labels)} uniform N=1024 instances x {args.c4_epochs} "
                               f"epochs; sweep ladder {lad[0]}..{lad[-1]}", "path": "c5"},
        "e_ref": float(np.asarray(ref["best_energy"])[0]),
        "time_to_99_s": {k: (v["wall_s"] if v else None) for k, v in hit.items()},
        "sweeps_to_99": {k: (v["sweeps"] if v else None) for k, v in hit.items()},
        "best_by_rung": {k: [r["mean_best"] for r in v[0]] for k, v in res.items()},
        "infer_s": t_inf, "marginal_confidence": float(np.mean(np.abs(p - 0.5))),
        "train_s": t_train, "label_s": t_labels,
        "loss_first_last": [float(np.mean(losses[:10])), float(np.mean(losses[-10:]))],
    }


def run_copy_calibration(args):
    """A plain device-to-device copy of known size (torch's copy kernel), whose
    FETCH_SIZE / WRITE_SIZE counters calibrate the traffic figures of the
    other kernels (bytes moved = 2 x size: one read, one write)."""
    import torch

    nbytes = 1 << 30
    x = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    y.copy_(x)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        y.copy_(x)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    return {"metric": "copy bandwidth", "value": 2 * nbytes * args.steps / el / 1e9, "unit": "GB/s",
            "n_gpus": 1, "steps": args.steps, "warmup": 1, "higher_is_better": True,
            "bytes_read_per_copy": nbytes, "bytes_written_per_copy": nbytes}


def run_tempering_sparse(args):
    """Replica exchange on the config-3 batch (512 sparse instances): 16
    ladders of 16 temperature slots (beta 0.1 .. 3.0 geometric) = 256
    replicas per instance, 2048 sweeps, an exchange round every --pt-interval
    sweeps, colour-class site order; plain annealing with the config-3
    schedule and budget runs in the same call for comparison."""
    import onesolver_amd
    from onesolver_amd import instances

    insts = instances.batch_c3(512)
    S, K, L, iv = 2048, 16, 16, args.pt_interval
    ts_slots = onesolver_amd.tscale_from_beta(np.geomspace(0.1, 3.0, K))
    ctx = onesolver_amd.Context(0)
    b = onesolver_amd.CsrBatch(ctx, insts)
    kw = dict(seed=1234, best_states=False, site_order=1)
    for _ in range(args.warmup):
        b.tempering(L, ts_slots, S, iv, **kw)
    t0 = time.perf_counter()
    best = np.inf
    for i in range(args.steps):
        out = b.tempering(L, ts_slots, S, iv, **kw)
        best = min(best, float(np.sum(out["inst_best_energy"])))
        log(f"step {i}: tempering sum of best energies {best:.4f}")
    elapsed = time.perf_counter() - t0
    ts_sa = onesolver_amd.tscale_from_beta(np.geomspace(0.1, 3.0, S))
    b.anneal(K * L, ts_sa, **kw)
    t1 = time.perf_counter()
    sa = b.anneal(K * L, ts_sa, **kw)
    sa_s = time.perf_counter() - t1
    b.close()
    ctx.close()
    sum_n = int(sum(len(d) for (_, _, _, d) in insts))
    attempts = float(K * L) * S * sum_n * args.steps
    return {
        "metric": "spin-flip attempts/s", "value": attempts / elapsed, "unit": "attempts/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp32 fields, fp64 energies",
        "data": "synthetic (seeded random d-regular graphs)",
        "config": {"workload": f"replica exchange, config-3 sparse batch (512 instances), {L} "
                               f"ladders x {K} slots (beta 0.1..3.0), {S} sweeps, exchange every "
                               f"{iv} sweeps, colour-class order",
                   "ladders": L, "slots": K, "sweeps": S, "exchange_interval": iv},
        "roofline": None,
        "sum_best_energy": best,
        "annealing_same_budget": {"sum_best_energy": float(np.sum(sa["inst_best_energy"])),
                                  "ms": sa_s * 1e3},
    }


def run_tempering(args):
    """Replica exchange on the config-2 instance (SK N=1024, seed 0): 256
    ladders of 64 temperature slots (beta 0.1 .. 3.0 geometric) = 16384
    replicas, 4096 sweeps, an exchange round every 16 sweeps, FP64 refresh
    every 512.  Plain annealing with the config-2 schedule and budget runs in
    the same call for comparison (its best energy and time)."""
    import onesolver_amd
    from onesolver_amd import instances

    n, S, K, L, iv = 1024, 4096, 64, 256, args.pt_interval
    q32, const = instances.sk_qubo(n, seed=0, instance=0)
    q = q32.astype(np.float64)
    ts_slots = onesolver_amd.tscale_from_beta(np.geomspace(0.1, 3.0, K))
    ctx = onesolver_amd.Context(0)
    prob = onesolver_amd.DenseProblem(ctx, q)
    kw = dict(seed=1234, refresh_interval=512, best_states=False)
    for _ in range(args.warmup):
        prob.tempering(L, ts_slots, S, iv, **kw)
    t0 = time.perf_counter()
    best, kernel_ms, launches = np.inf, 0.0, 0
    for i in range(args.steps):
        out = prob.tempering(L, ts_slots, S, iv, **kw)
        best = min(best, out["best_energy"])
        kernel_ms += out["stats"]["kernel_ms"]
        launches += out["stats"]["kernel_launches"]
        log(f"step {i}: tempering best {out['best_energy']:.4f}")
    elapsed = time.perf_counter() - t0
    # plain annealing, same instance, replicas and sweeps (config 2)
    ts_sa = onesolver_amd.tscale_from_beta(np.geomspace(0.1, 3.0, S))
    t1 = time.perf_counter()
    sa = prob.anneal(K * L, ts_sa, **kw)
    sa_s = time.perf_counter() - t1
    prob.close()
    ctx.close()
    attempts = float(K * L) * S * n * args.steps
    avg_launch_s = kernel_ms / launches / 1e3
    flops_per_launch = 4.0 * n * (K * L * (S / (launches / args.steps)) * n)
    achieved = flops_per_launch / avg_launch_s / 1e12
    return {
        "metric": "spin-flip attempts/s", "value": attempts / elapsed, "unit": "attempts/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "bf16 hi+lo MFMA, fp32 fields, fp64 energies",
        "data": "synthetic (Philox-generated SK instance, seed 0)",
        "config": {"workload": f"replica exchange, dense SK QUBO N=1024, {L} ladders x {K} "
                               f"slots (beta 0.1..3.0), {S} sweeps, exchange every {iv} sweeps",
                   "n": n, "ladders": L, "slots": K, "sweeps": S, "exchange_interval": iv},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": PEAK_BF16_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / PEAK_BF16_TFLOPS, "traffic": None,
                     "kernel": "dense_sweep_kernel<8>", "avg_launch_ms": avg_launch_s * 1e3,
                     "algorithmic_flops_per_launch": flops_per_launch},
        "best_energy": best, "best_energy_ising": best + const,
        "annealing_same_budget": {"best_energy": sa["best_energy"],
                                  "best_energy_ising": sa["best_energy"] + const,
                                  "ms": sa_s * 1e3},
    }


def run_population(args):
    """Population annealing at the annealing configs' budgets, beside plain
    annealing on the same schedule (best energies and time).  Dense: the
    config-2 instance (SK N=1024, seed 0), 16384 replicas, 64 temperature
    steps (beta 0.1 .. 3.0 geometric) x 64 sweeps = 4096 sweeps, FP64 refresh
    every 512.  Sparse: the config-3 batch (512 instances), 256 replicas,
    32 steps x 64 sweeps = 2048 sweeps, colour-class site order."""
    import onesolver_amd
    from onesolver_amd import instances

    K, theta = (64, 64) if args.pa_path == "dense" else (32, 64)
    S = K * theta
    betas = np.geomspace(0.1, 3.0, K)
    ts_steps = onesolver_amd.tscale_from_beta(betas)
    ts_sa = np.repeat(ts_steps, theta)
    ctx = onesolver_amd.Context(0)
    if args.pa_path == "dense":
        n, R = 1024, 16384
        q32, const = instances.sk_qubo(n, seed=0, instance=0)
        prob = onesolver_amd.DenseProblem(ctx, q32.astype(np.float64))
        kw = dict(seed=1234, refresh_interval=512, best_states=False)
        run = lambda: prob.population(R, ts_steps, theta, **kw)  # noqa: E731
        plain = lambda: prob.anneal(R, ts_sa, **kw)  # noqa: E731
        best_of = lambda out: float(out["best_energy"])  # noqa: E731
        sum_n = n
        work = f"dense SK QUBO N=1024, {R} replicas"
    else:
        insts = instances.batch_c3(512)
        R = 256
        prob = onesolver_amd.CsrBatch(ctx, insts)
        kw = dict(seed=1234, best_states=False, site_order=1)
        run = lambda: prob.population(R, ts_steps, theta, **kw)  # noqa: E731
        plain = lambda: prob.anneal(R, ts_sa, **kw)  # noqa: E731
        best_of = lambda out: float(np.sum(out["inst_best_energy"]))  # noqa: E731
        sum_n = int(sum(len(d) for (_, _, _, d) in insts))
        work = f"config-3 sparse batch (512 instances), {R} replicas, colour-class order"
        const = 0.0
    for _ in range(args.warmup):
        run()
    t0 = time.perf_counter()
    best = np.inf
    for i in range(args.steps):
        out = run()
        best = min(best, best_of(out))
        log(f"step {i}: population best {best_of(out):.4f}")
    elapsed = time.perf_counter() - t0
    plain()
    t1 = time.perf_counter()
    sa = plain()
    sa_s = time.perf_counter() - t1
    prob.close()
    ctx.close()
    attempts = float(R) * S * sum_n * args.steps
    return {
        "metric": "spin-flip attempts/s", "value": attempts / elapsed, "unit": "attempts/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp32 fields, fp64 energies",
        "data": "synthetic (Philox-generated instances)",
        "config": {"workload": f"population annealing, {work}, {K} steps (beta 0.1..3.0) x "
                               f"{theta} sweeps, resampling between steps",
                   "path": args.pa_path, "replicas": R, "steps_K": K, "sweeps_per_step": theta,
                   "sweeps": S},
        "roofline": None,
        "best_energy": best, "best_energy_ising": best + const,
        "annealing_same_budget": {"best_energy": best_of(sa),
                                  "best_energy_ising": best_of(sa) + const,
                                  "ms": sa_s * 1e3},
    }


def run_exhaustive(args):
    """All 2^n states of a dense U[-1,1] instance (os_exhaustive), n=32 by
    default.  One step = one search, Q uploaded by the call (0.01 ms)."""
    import onesolver_amd
    n = args.exh_n
    rng = np.random.default_rng(0)
    q = np.triu(rng.uniform(-1.0, 1.0, (n, n)))
    ctx = onesolver_amd.Context(0)
    for _ in range(args.warmup):
        ctx.exhaustive(q)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        st, e = ctx.exhaustive(q)
    elapsed = time.perf_counter() - t0
    ctx.close()
    states = float(2 ** n) * args.steps
    # per state: one FP64 add for the energy and kLow - 1 = 4 for the low
    # fields of the Gray-code walk (the second pass re-walks only the threads
    # whose minimum is within the re-ranking tolerance)
    flops = 5.0 * states
    achieved = flops / elapsed / 1e12
    res = {
        "metric": "exhaustive states/s", "value": states / elapsed, "unit": "states/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp64", "data": "synthetic (numpy seed 0, U[-1,1] upper triangle)",
        "config": {"workload": f"exhaustive search, dense n={n}, all 2^{n} states, lowest index "
                               f"on ties", "n": n},
        "roofline": {"bound": "fp64", "achieved": achieved, "peak": PEAK_FP64_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / PEAK_FP64_TFLOPS, "traffic": None,
                     "kernel": "exh_gray_kernel, both passes (whole os_exhaustive call)",
                     "avg_launch_ms": elapsed / args.steps * 1e3,
                     "algorithmic_flops_per_launch": 5.0 * 2 ** n},
        "best_energy": e,
    }
    if not args.no_cpu_baseline:
        import oracle
        threads = _host_threads()
        m = 31  # a full search of 2^31 states, Gray-code restatement on all cores
        qs = np.triu(q[:m, :m]) + np.triu(q[:m, :m], 1).T
        t0 = time.perf_counter()
        oracle.exhaustive(qs, m, threads=threads)
        port_s = time.perf_counter() - t0
        ml = 22  # the reference's literal O(N^2) per-state loop on one thread
        t0 = time.perf_counter()
        oracle.exhaustive(qs[:ml, :ml], ml, literal=True)
        lit_s = time.perf_counter() - t0
        res["cpu_baseline"] = {
            "value": 2.0 ** m / port_s, "unit": "states/s", "cores": threads, "kind": "port",
            "sample": f"Gray-code restatement with literal re-ranking, all 2^{m} states of the "
                      f"leading {m}x{m} block, {port_s:.1f} s",
            "literal_one_thread": {"value": 2.0 ** ml / lit_s, "unit": "states/s", "cores": 1,
                                   "sample": f"reference per-state O(N^2) loop, n={ml}, "
                                             f"{lit_s:.2f} s"}}
    return res


def run_batch_cli(args):
    """one-solver-batch over the config-3 batch written as 512 `.qubo` files:
    GPU parse of every file, one batched sparse anneal (256 replicas, 2048
    sweeps, beta 0.1 -> 3.0 as inverse temperature), one Solution CSV per
    instance.  One step = one run of the CLI process, files to CSVs."""
    import subprocess
    import tempfile

    import onesolver_amd
    from onesolver_amd import instances
    insts = instances.batch_c3(512)
    R, S = 256, args.sweeps_c3
    tmp = args.batch_dir or tempfile.mkdtemp(prefix="osbatch_")
    inp, out = os.path.join(tmp, "in"), os.path.join(tmp, "out")
    os.makedirs(inp, exist_ok=True)
    os.makedirs(out, exist_ok=True)
    nbytes = 0
    for k, (rp, cl, vl, dg) in enumerate(insts):
        n = len(dg)
        rows = np.repeat(np.arange(n), np.diff(rp))
        up = cl > rows
        lines = ["p qubo 0 %d %d %d\n" % (n, n, int(up.sum()))]
        lines += ["%d %d %.17g\n" % (i, i, dg[i]) for i in range(n)]
        lines += ["%d %d %.17g\n" % t for t in zip(rows[up], cl[up], vl[up])]
        text = "".join(lines)
        nbytes += len(text)
        with open(os.path.join(inp, "c3_%03d.qubo" % k), "w") as f:
            f.write(text)
    cmd = [onesolver_amd.BATCH_CLI_PATH, "--input", inp, "--output", out, "--num-iter", str(S),
           "--num-tries", str(R), "--beta-min", "0.1", "--beta-max", "3", "--beta-kind",
           "inverse", "--path", "sparse", "--site-order", "colour"]
    for _ in range(args.warmup):
        subprocess.run(cmd, check=True, capture_output=True, timeout=300)
    phases = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        r = subprocess.run(cmd, check=True, capture_output=True, text=True, timeout=300)
        ph = {}
        for line in r.stdout.splitlines():
            for key in ("Read and parse", "Anneal", "Write"):
                if line.startswith(key + ": "):
                    ph[key] = float(line.split(": ")[1].split()[0])
        phases.append(ph)
    elapsed = time.perf_counter() - t0
    csvs = len([f for f in os.listdir(out) if f.endswith(".csv")])
    assert csvs == len(insts), csvs
    ns = np.array([len(d) for (_, _, _, d) in insts], dtype=np.float64)
    attempts = float(ns.sum()) * R * S * args.steps
    mean = {k: float(np.mean([p[k] for p in phases])) for k in phases[0]}
    res = {
        "metric": "instances/s", "value": len(insts) * args.steps / elapsed, "unit": "instances/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp32 fields, fp64 energies",
        "data": "synthetic (seeded random d-regular graphs, written as .qubo text)",
        "config": {"workload": f"one-solver-batch: 512 .qubo files ({nbytes / 1e6:.1f} MB) -> "
                               f"512 CSVs, sparse path (colour-class order), {R} replicas, "
                               f"{S} sweeps",
                   "instances": len(insts), "replicas": R, "sweeps": S,
                   "sum_n": int(ns.sum()), "text_bytes": nbytes},
        "attempts_per_s": attempts / elapsed,
        "phase_ms": mean,
        "roofline": None,
    }
    if not args.no_cpu_baseline:
        imax = int(np.argmax(ns))
        cb = cpu_baseline_csr(insts[imax], onesolver_amd.tscale_from_beta(
            np.geomspace(0.1, 3.0, S)))
        # the same attempt rate over the batch's mean instance: instances per second
        per_inst = float(ns.mean()) * R * S
        res["cpu_baseline"] = {"value": cb["value"] / per_inst, "unit": "instances/s",
                               "cores": cb["cores"], "kind": cb["kind"],
                               "sample": cb["sample"] + f"; {cb['value']:.3g} attempts/s over "
                                         f"{per_inst:.3g} attempts per mean instance"}
    return res


def _dense_qubo_text(n, seed=0, chunk=1 << 22):
    """`.qubo` text of a dense n-variable instance: the header, then the
    upper triangle row by row as fixed-width lines "iiii jjjj +0.ddddddddddddddddd"
    (indices right-aligned in blanks, 17 random fraction digits, random sign).
    Built with numpy, without per-line Python formatting."""
    rng = np.random.default_rng(seed)
    w = len(str(n - 1))
    iu, ju = np.triu_indices(n)
    L = len(iu)
    W = 2 * w + 23
    head = ("p qubo 0 %d %d %d\n" % (n, n, L - n)).encode()
    out = np.empty(len(head) + L * W, np.uint8)
    out[:len(head)] = np.frombuffer(head, np.uint8)
    body = out[len(head):].reshape(L, W)

    def field(vals, cols):
        v = vals.astype(np.int64)
        for c in range(w - 1, -1, -1):
            d = (v % 10).astype(np.uint8) + 48
            if c < w - 1:
                d = np.where(v == 0, np.uint8(32), d)
            cols[:, c] = d
            v //= 10

    for a in range(0, L, chunk):
        b = min(L, a + chunk)
        blk = body[a:b]
        field(iu[a:b], blk[:, :w])
        blk[:, w] = 32
        field(ju[a:b], blk[:, w + 1:2 * w + 1])
        blk[:, 2 * w + 1] = 32
        blk[:, 2 * w + 2] = np.where(rng.integers(0, 2, b - a) == 1, 43, 45).astype(np.uint8)
        blk[:, 2 * w + 3] = 48
        blk[:, 2 * w + 4] = 46
        blk[:, 2 * w + 5:2 * w + 22] = rng.integers(48, 58, (b - a, 17), dtype=np.uint8)
        blk[:, 2 * w + 22] = 10
    return out, L


def run_ingest(args):
    """`.qubo` text ingest on the GPU (os_qubo_parse_device): the dense
    N=8192 instance as text (33.6M coefficient lines, 1.04 GB), already in
    GPU memory.  One step = one parse: line index, per-line parse, first-wins
    dedupe and compaction into the term arrays, host checks included."""
    import onesolver_amd
    n = args.ingest_n
    t0 = time.perf_counter()
    text, L = _dense_qubo_text(n)
    log(f"ingest text n={n}: {len(text) / 1e9:.3f} GB, {L} lines "
        f"({time.perf_counter() - t0:.1f} s to build)")
    ctx = onesolver_amd.Context(0)
    dev = onesolver_amd.DeviceBytes(ctx, text)
    src = dev.source
    for _ in range(args.warmup):
        onesolver_amd.Qubo(ctx, device=src).close()
    times = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        q = onesolver_amd.Qubo(ctx, device=src)
        times.append(q.parse_times())
        info = (q.n, q.num_linear, q.num_quadratic)
        q.close()
    elapsed = time.perf_counter() - t0
    assert info == (n, n, L - n), info
    ms = np.mean(np.asarray(times, np.float64), axis=0)
    # host text: the same parse with the text handed over from host memory
    # (upload through pinned staging buffers included)
    host_text = text.tobytes()
    onesolver_amd.Qubo(ctx, host_text).close()
    th = time.perf_counter()
    onesolver_amd.Qubo(ctx, host_text).close()
    host_s = time.perf_counter() - th
    del host_text
    dev.close()
    ctx.close()
    nbytes = len(text)
    # parse_lines: reads the text and the line starts (8 B), writes kind, i,
    # j and the value (17 B) per line
    alg = nbytes + 25.0 * (L + 1)
    achieved = alg / (ms[1] * 1e-3) / 1e9
    res = {
        "metric": "ingest bytes/s", "value": nbytes * args.steps / elapsed, "unit": "B/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "u8 text -> fp64",
        "data": "synthetic (numpy seed 0: fixed-width lines, 17 random fraction digits)",
        "config": {"workload": f".qubo text of a dense N={n} instance ({L} coefficient lines, "
                               f"{nbytes / 1e9:.2f} GB) already in GPU memory -> distinct terms",
                   "n": n, "lines": L, "bytes": nbytes},
        "lines_per_s": L * args.steps / elapsed,
        "gpu_ms": {"line_index": float(ms[0]), "parse_lines": float(ms[1]),
                   "dedupe_compact": float(ms[2])},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": achieved / PEAK_HBM_GBS,
                     "traffic": TRAFFIC_INGEST_8192 if n == 8192 else None,
                     "kernel": "parse_lines_kernel<true> (no round here is longer than its "
                               "16 KiB LDS span)", "avg_launch_ms": float(ms[1]),
                     "algorithmic_bytes_per_launch": alg},
        "host_text_bytes_per_s": nbytes / host_s,
    }
    if not args.no_cpu_baseline:
        import oracle
        # the first m lines of the same text, under a header that declares them
        m = min(L, 2_000_000)
        W = 2 * len(str(n - 1)) + 23
        head = len(text) - L * W
        iu, ju = np.triu_indices(n)
        nl = int(np.count_nonzero(iu[:m] == ju[:m]))
        sample = (("p qubo 0 %d %d %d\n" % (n, nl, m - nl)).encode() +
                  text[head:head + m * W].tobytes())
        t0 = time.perf_counter()
        code, msg, _ = oracle.qubo_parse(sample)
        cpu_s = time.perf_counter() - t0
        assert code == 0, msg
        res["cpu_baseline"] = {
            "value": len(sample) / cpu_s, "unit": "B/s", "cores": 1, "kind": "port",
            "sample": f"the reference grammar restated (sequential reader, strtod, first-wins "
                      f"hash), first {m} lines ({len(sample) / 1e6:.0f} MB), {cpu_s:.2f} s"}
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="c2",
                    choices=["c2", "c1", "c3", "c4", "c5", "exh", "pt", "pa", "ingest", "batch",
                             "copy"])
    ap.add_argument("--pa-path", choices=["dense", "sparse"], default="dense")
    ap.add_argument("--ingest-n", type=int, default=8192)
    ap.add_argument("--pt-interval", type=int, default=16)
    ap.add_argument("--pt-path", choices=["dense", "sparse"], default="dense")
    ap.add_argument("--exh-n", type=int, default=32)
    ap.add_argument("--c5-replicas", type=int, default=4096)
    ap.add_argument("--c5-sweeps", type=int, default=1024)
    ap.add_argument("--c5-group", type=int, default=0,
                    help="blocks of 64 spins per field pass (0: the library's choice)")
    ap.add_argument("--c5-chunk", type=int, default=0,
                    help="pass launch: blocks whose operands load per round trip (0: default)")
    ap.add_argument("--c5-wgs", type=int, default=0,
                    help="pass launch: persistent work-groups per CU (0: default)")
    ap.add_argument("--c4-train", type=int, default=4096)
    ap.add_argument("--c4-epochs", type=int, default=2)
    ap.add_argument("--c4-eval", type=int, default=32)
    ap.add_argument("--c4-path", choices=["dense", "sparse", "c5"], default="dense",
                    help="sparse: the warm start on the config-3 (mixed-sparse) batch; c5: "
                         "learned vs plain init on the config-5 instance")
    ap.add_argument("--sweeps-c3", type=int, default=2048)
    ap.add_argument("--c3-order", choices=["colour", "sequential"], default="colour")
    ap.add_argument("--batch-dir", default="",
                    help="batch workload: write the .qubo files to DIR/in and the CSVs to "
                         "DIR/out and keep them (default: a temporary directory)")
    ap.add_argument("--c3-no-tune", action="store_true",
                    help="skip the launch-shape tuner (library default shape)")
    ap.add_argument("--n", type=int, default=1024)
    ap.add_argument("--replicas", type=int, default=16384)
    ap.add_argument("--sweeps", type=int, default=4096)
    ap.add_argument("--refresh", type=int, default=512)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build runs on one GPU; multi-GPU runs are not part of it "
              "(replicas would shard across GPUs with one final min-reduction)", file=sys.stderr)
        sys.exit(2)
    if args.workload == "copy":
        print(json.dumps(run_copy_calibration(args)))
        return
    if args.workload == "c4":
        run = {"sparse": run_warmstart_c4_sparse, "c5": run_warmstart_c5}.get(
            args.c4_path, run_warmstart_c4)
        print(json.dumps(run(args)))
        return
    if args.workload == "batch":
        print(json.dumps(run_batch_cli(args)))
        return
    if args.workload == "ingest":
        print(json.dumps(run_ingest(args)))
        return
    if args.workload == "exh":
        print(json.dumps(run_exhaustive(args)))
        return
    if args.workload == "pt":
        print(json.dumps(run_tempering_sparse(args) if args.pt_path == "sparse"
                         else run_tempering(args)))
        return
    if args.workload == "pa":
        print(json.dumps(run_population(args)))
        return
    if args.workload == "c5":
        res, _ = run_dense_c5(args)
        if not args.no_cpu_baseline:
            from onesolver_amd import instances
            log("timing the CPU baseline (bounded sample)")
            q = instances.dense_uniform(8192, seed=0, instance=0).astype(np.float64)
            res["cpu_baseline"] = cpu_baseline_dense(q, 8192)
        print(json.dumps(res))
        return
    if args.workload == "c3":
        res, (inst, ts) = run_csr_c3(args)
        if not args.no_cpu_baseline:
            log("timing the CPU baseline (bounded sample)")
            res["cpu_baseline"] = cpu_baseline_csr(inst, ts)
        print(json.dumps(res))
        return
    if args.workload == "c1":
        res, q = run_reference_c1(args)
    else:
        res, q = run_dense(args)
    if not args.no_cpu_baseline:
        log("timing the CPU baseline (bounded sample)")
        res["cpu_baseline"] = cpu_baseline_dense(q, q.shape[0])
    print(json.dumps(res))


if __name__ == "__main__":
    main()
