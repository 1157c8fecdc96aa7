/*
 * onesolver.h -- C ABI of the MI355X simulated-annealing QUBO library
 * (libonesolver.so).
 *
 * The reference (oneSolver) is a header-only C++ template library with no FFI
 * layer; each entry point below names the reference interface it replaces.
 * Conventions:
 *   - every call returns OS_OK (0) or a negative OS_ERR_* code; the message of
 *     the last failing call on the calling thread is os_last_error();
 *   - host arrays are caller-owned and only read/written during the call
 *     (copied to and from device memory inside);
 *   - a context owns one GPU, one HIP stream and its device buffers; calls on
 *     one context are synchronous and must not overlap; separate contexts
 *     are independent;
 *   - QUBO matrices are dense row-major symmetric N x N ("flat" in the
 *     reference: include/helpers/qubo_helpers.hpp:26-44) with the linear
 *     terms on the diagonal; energy E(x) = sum_{i<=j} Q_ij x_i x_j
 *     (include/simulated_annealing/annealing.hpp:31-40).
 */
#ifndef ONESOLVER_H_
#define ONESOLVER_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OS_OK 0
#define OS_ERR_INVALID (-1)     /* invalid argument (reference: std::invalid_argument) */
#define OS_ERR_HIP (-2)         /* HIP runtime error (reference: sycl::exception) */
#define OS_ERR_UNSUPPORTED (-3) /* size or option outside what the kernels support */
#define OS_ERR_INTERNAL (-4)

typedef struct os_ctx os_ctx;

/* Library version as major*10000 + minor*100 + patch. */
int os_version(void);

/* Message of the last failing call on this thread ("" if none). */
const char *os_last_error(void);

/* Number of visible GPUs.  Replaces the SYCL device enumeration behind
 * devices::construct_device_selector (include/helpers/devices.hpp:27-42). */
int os_device_count(int *count);

/* Create / destroy a context on GPU `device`.  Replaces the sycl::queue of
 * app/one-solver-anneal.cpp:139-147.  Objects created on a context (dense
 * problems and batches, sparse batches, .qubo models, networks) use its
 * stream until they are destroyed: destroy them first. */
int os_ctx_create(int device, os_ctx **out);
int os_ctx_destroy(os_ctx *ctx);

/* Device name ("Using device: ..." line, app/one-solver-anneal.cpp:149-151). */
int os_ctx_device_name(os_ctx *ctx, char *buf, size_t len);

/* Reference-semantics simulated annealing.  Replaces sa::anneal
 * (include/simulated_annealing/annealing.hpp:55-140):
 *   - num_tries independent replicas, each starting from uniform random bits;
 *   - num_iter * sweeps_per_beta attempts per replica; attempt t proposes a
 *     uniformly random site and uses beta[t / sweeps_per_beta];
 *   - accept iff E_new < E_cur || exp((E_cur - E_new) / beta) > u, u~U[0,1)
 *     (beta acts as a temperature, as in the reference);
 *   - the best state of each replica is tracked after every attempt.
 * q_sym: N x N FP64.  beta: num_iter values.  The RNG is Philox4x32-10 keyed
 * by `seed` (the reference's fixed seed is 1234).
 * Outputs (each may be NULL): best_states [num_tries x n] (0/1 bytes),
 * best_energies [num_tries] (FP64, recomputed from the returned states),
 * best_state [n] and best_energy: the replica with the lowest energy, first
 * index on ties (annealing.hpp:134-139). */
int os_anneal_reference(os_ctx *ctx, int n, const double *q_sym, const double *beta,
                        int num_iter, int sweeps_per_beta, uint32_t num_tries, uint64_t seed,
                        uint8_t *best_states, double *best_energies, uint8_t *best_state,
                        double *best_energy);

/* Bytes of the fixed per-GPU scratch array of os_anneal_reference_noalloc. */
#define OS_NOALLOC_SCRATCH_BYTES (16u << 20)

/* os_anneal_reference for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  No counterpart in the reference.  The same chains, energies
 * and tie rule as os_anneal_reference, with every array in device memory of
 * the context's GPU: d_q_sym (n x n FP64) and d_beta (num_iter FP64) are read;
 * the outputs, each optional, are d_best_states [num_tries x n] (0/1 bytes),
 * d_best_energies [num_tries] (FP64, recomputed from the states) and
 * d_best_index [1] (int32: the replica of the lowest energy, first index on
 * ties).  Everything else lives in a fixed scratch array of
 * OS_NOALLOC_SCRATCH_BYTES per GPU, held until the process ends: the call
 * makes no hipMalloc or hipFree.  It returns OS_ERR_UNSUPPORTED when the run
 * needs more scratch than that: 4*ceil(n/32) + 16 + 8*ceil(n/64) bytes per
 * try (each part rounded up to 256 bytes), plus 8*n per try for n > 4096.
 * All calls on one GPU share its scratch, so calls that overlap on one GPU,
 * from any context, race.  Synchronous, like the other entry points. */
int os_anneal_reference_noalloc(os_ctx *ctx, int n, const double *d_q_sym, const double *d_beta,
                                int num_iter, int sweeps_per_beta, uint32_t num_tries,
                                uint64_t seed, uint8_t *d_best_states, double *d_best_energies,
                                int32_t *d_best_index);
/* OS_NOALLOC_SCRATCH_BYTES (returns the size, not a status). */
size_t os_noalloc_scratch_bytes(void);

/* FP64 QUBO energies of `count` states (0/1 bytes, count x n).  Replaces
 * sa::energy (include/simulated_annealing/annealing.hpp:31-40). */
int os_energy_fp64(os_ctx *ctx, int n, const double *q_sym, uint32_t count,
                   const uint8_t *states, double *energies);

/* Exact minimum by enumeration of all 2^n states.  Replaces exhaustive::solve
 * (include/exhaustive/exhaustive.hpp:29-167): E(x) = sum_{i <= j} q[i][j]
 * x_i x_j over the upper triangle of q_upper (n x n FP64, row-major; the
 * diagonal holds the linear terms, entries below it are ignored), summed in
 * the reference's order (i ascending, then j >= i ascending); of equal
 * energies the lowest state index wins, bit i of the index being variable i.
 * n <= 40 (the reference: n <= 30, exhaustive.hpp:67; 32 for its state
 * conversion, ulong_to_vec.hpp:23-26).  Non-finite coefficients are supported
 * up to n = 30; a state whose energy is NaN or +inf never wins, as in
 * exhaustive.hpp:108-132.  Outputs: best_state [n] (0/1 bytes) and
 * best_energy. */
int os_exhaustive(os_ctx *ctx, int n, const double *q_upper, uint8_t *best_state,
                  double *best_energy);

/* ---- sweep-mode annealing (configs quoted in "sweeps") -------------------
 * A sweep visits every site once in the fixed order j = 0..N-1; site j of
 * sweep t flips iff dE < -ln(u) * tscale[t] with u the open-interval 24-bit
 * uniform of Philox counter (j>>2, replica, t, (instance<<8)|0x11), i.e. with
 * probability min(1, exp(-dE / tscale[t])).  tscale[t] = 1/beta_t for an
 * inverse-temperature schedule; tscale[t] = beta_t reproduces the reference's
 * use of beta as a temperature (annealing.hpp:107).  The best state of each
 * replica is tracked at sweep boundaries; returned energies are FP64
 * recomputations of the returned states.  No reference counterpart: the
 * reference only has random-site attempts (annealing.hpp:97-101). */
typedef struct os_sweep_params {
  uint32_t num_replicas;       /* replicas (rounded up to a multiple of 64 inside) */
  int32_t num_sweeps;          /* sweeps per replica */
  const float *tscale;         /* [num_sweeps] host array */
  uint64_t seed;               /* Philox key */
  uint32_t instance;           /* stream id of this instance (< 2^24) */
  int32_t refresh_interval;    /* sweeps between FP64 field refreshes; <= 0: once at the end */
  const uint8_t *init_states;  /* optional [num_replicas x n] start states (0/1) */
  const float *init_marginals; /* optional [n] warm start: x_j ~ Bernoulli(p_j) */
} os_sweep_params;

typedef struct os_sweep_stats {
  double kernel_ms;        /* summed duration of the sweep kernels (HIP events) */
  double total_ms;         /* the whole call on the device stream */
  int32_t kernel_launches; /* sweep-kernel launches */
  double max_field_drift;  /* max |h_fp32 - h_fp64| over all refreshes */
  uint64_t attempts;       /* num_replicas x num_sweeps x n spin-flip attempts */
  double pass_avg_ms;      /* dense n > 1024: average duration of one field pass
                              (HIP events around a sample of the passes); else 0 */
  int32_t pass_launches;   /* dense n > 1024: field passes; else 0 */
} os_sweep_stats;

/* Replica exchange (parallel tempering) on a dense batch, n <= 1024.  Each
 * instance runs num_ladders ladders of num_temperatures slots; replica
 * r = l * num_temperatures + k starts in slot k of ladder l.  A replica sweeps
 * at the temperature scale (1/beta) of its slot, with the thresholds of
 * sweep-mode annealing.  After every exchange_interval sweeps, round t pairs
 * the slots (k, k+1), k = t mod 2, t mod 2 + 2, ..., of every ladder; the two
 * replicas swap slots with probability min(1, exp((beta_k - beta_{k+1})
 * (E_k - E_{k+1}))), drawn as u < exp(.) with u = (w >> 8) 2^-24 and w word 0
 * of Philox counter (k, l, t, ((instance + i) << 8) | 0x13).  The energies
 * are the running energies of the launch (FP64 after a refresh).  No
 * reference counterpart: the reference recommends replica exchange
 * (benchmarks/annealing/performance.md:52-59) but anneals only. */
typedef struct os_tempering_params {
  uint32_t num_ladders;
  uint32_t num_temperatures; /* >= 2 slots per ladder */
  const float *tscale;       /* [num_temperatures] host array, > 0 */
  int32_t num_sweeps;
  int32_t exchange_interval; /* sweeps between exchange rounds, >= 1 */
  uint64_t seed;             /* Philox key */
  uint32_t instance;         /* stream id of this instance (< 2^24) */
  int32_t refresh_interval;  /* a multiple of exchange_interval, or <= 0 (end only) */
} os_tempering_params;

/* Population annealing on a dense batch.  Each instance anneals a population
 * of num_replicas sweep-mode replicas through num_steps temperature scales:
 * step k runs sweeps_per_step sweeps at tscale[k] (sweep counters
 * k * sweeps_per_step, ...).  After every step but the last, the population
 * is resampled systematically: with E the running energies and
 * dbeta = 1/tscale[k+1] - 1/tscale[k], replica j continues from the state,
 * fields and energy of replica anc[j], where W_i is the running sum of
 * exp(-dbeta (E_i - min E)) and anc[j] is the first i with
 * W_i > (j + u) / R * W_{R-1}; u is the 53-bit uniform of words 0, 1 of
 * Philox counter (k, 0, 0, ((instance + i) << 8) | 0x14).  The best state of
 * each replica index stays with the index.  No reference counterpart: the
 * reference recommends population annealing
 * (benchmarks/annealing/performance.md:52-59) but anneals only. */
typedef struct os_population_params {
  uint32_t num_replicas;    /* population size R per instance */
  int32_t num_steps;        /* K >= 1 temperature steps */
  int32_t sweeps_per_step;  /* >= 1 */
  const float *tscale;      /* [num_steps] host array, > 0 */
  uint64_t seed;            /* Philox key */
  uint32_t instance;        /* stream id of this instance (< 2^24) */
  int32_t refresh_interval; /* a multiple of sweeps_per_step, or <= 0 (end only) */
} os_population_params;

typedef struct os_dense os_dense;

/* Upload a dense QUBO (q_sym: n x n FP64 symmetric, the reference's flat
 * matrix) for sweep-mode annealing; keeps Q as FP64 and as bf16 hi+lo pairs
 * on the GPU.  n <= 1024: fields in registers, one launch per segment;
 * 1024 < n <= 16384: fields in HBM, blocks of 64 spins (os_sweep_stats
 * reports the field passes). */
int os_dense_create(os_ctx *ctx, int n, const double *q_sym, os_dense **out);
/* A batch of num_instances same-size dense QUBOs (q_sym: num_instances
 * consecutive n x n matrices), annealed together in one launch per segment;
 * instance i draws its random numbers as stream instance + i.  No reference
 * counterpart: the reference anneals one model per sa::anneal call
 * (include/simulated_annealing/annealing.hpp:55-58). */
int os_dense_batch_create(os_ctx *ctx, int num_instances, int n, const double *q_sym,
                          os_dense **out);
/* A batch of num_instances SK instances of the dense benchmark distribution
 * (J_ij ~ N(0,1)/sqrt(n) fp32, h = 0, as a QUBO: Q_ij = 4 J_ij, Q_ii =
 * -2 sum_j J_ij), generated on the GPU from Philox key (seed, instance_base +
 * i): no host generation or transfer.  No reference counterpart. */
int os_dense_batch_create_sk(os_ctx *ctx, int num_instances, int n, uint64_t seed,
                             uint32_t instance_base, os_dense **out);
/* Dense uniform instances (Q_ij, i <= j, i.i.d. U[-1,1) fp32, mirrored),
 * generated on the GPU from Philox key (seed, instance_base + i).  No
 * reference counterpart. */
int os_dense_batch_create_uniform(os_ctx *ctx, int num_instances, int n, uint64_t seed,
                                  uint32_t instance_base, os_dense **out);
/* Copy the FP64 matrices of a dense batch to the host ([I x n x n]). */
int os_dense_copy_q(os_dense *dense, double *q_out);
int os_dense_destroy(os_dense *dense);

/* Run num_replicas sweep-mode chains per instance.  Outputs (each may be
 * NULL), instance-major for a batch of I instances (I = 1 for os_dense_create):
 * best_energies [I x R], best_states [I x R x n], final_states [I x R x n],
 * best_state [I x n] + best_energy [I] (lowest replica, first index on ties).
 * init_states is [I x R x n], init_marginals [I x n]. */
int os_dense_anneal(os_dense *dense, const os_sweep_params *params, double *best_energies,
                    uint8_t *best_states, uint8_t *final_states, uint8_t *best_state,
                    double *best_energy, os_sweep_stats *stats);
/* Autotuner of the streaming path (n > 1024).  Times a probe run of
 * min(num_sweeps, 4) sweeps (the params' replicas, schedule and seed, no
 * outputs) for every candidate: blocks per field pass (3, 4, 5, 6, 8) x
 * persistent pass work-groups per CU (1, 2), keeps the fastest on the handle
 * for its later runs and returns it (best_ms: its summed sweep-kernel time).
 * Neither knob changes results: every field is the same fp32 sum in the same
 * block order.  The OS_STREAM_GROUP / OS_STREAM_PASS_WGS_PER_CU environment
 * variables still take precedence.  For n <= 1024 nothing is tuned and the
 * outputs are 0: the register-resident kernel fixes k = 16 and 64 replicas
 * per work-group by its register layout.  No reference counterpart. */
int os_dense_tune(os_dense *dense, const os_sweep_params *params, int32_t *group,
                  int32_t *pass_wgs_per_cu, double *best_ms);
/* Replica exchange on `dense` (see os_tempering_params).  Outputs as in
 * os_dense_anneal, indexed by replica; final_slots (optional, [I][R]) holds
 * the slot of every replica at the end. */
int os_dense_tempering(os_dense *dense, const os_tempering_params *params, double *best_energies,
                       uint8_t *best_states, uint8_t *final_states, uint32_t *final_slots,
                       uint8_t *best_state, double *best_energy, os_sweep_stats *stats);
/* Population annealing on `dense` (see os_population_params).  Outputs as in
 * os_dense_anneal, indexed by replica; ancestors (optional,
 * [I][num_steps - 1][R]) holds anc of every resampling round. */
int os_dense_population(os_dense *dense, const os_population_params *params,
                        double *best_energies, uint8_t *best_states, uint8_t *final_states,
                        uint32_t *ancestors, uint8_t *best_state, double *best_energy,
                        os_sweep_stats *stats);

/* ---- learned warm start --------------------------------------------------
 * A 3-layer message-passing network (hidden 64, bf16 weights, fp32
 * accumulation) over the coupling graph of dense instances predicts per-spin
 * marginals p_j of good annealed states; os_dense_anneal's init_marginals
 * seeds replicas from them.  Trained on the GPU from best states of the plain
 * annealer.  Labels are 0/1 states [count x n]; a state and its complement
 * have the same SK energy, so labels are taken with spin 0 flipped to 1 (and
 * the network pins spin 0).  Parameters are exchanged as fp32 in the packed
 * order: per layer Ws [fin x 64], Wm [fin x 64], b [64] (fin = 4, 64, 64),
 * then wo [64], bo [1].  No reference counterpart: the reference only starts
 * from uniform random bits (include/simulated_annealing/annealing.hpp:88-91). */
typedef struct os_mpnn os_mpnn;

int os_mpnn_num_params(void); /* returns the count (17153), not a status */
int os_mpnn_create(os_ctx *ctx, uint64_t seed, os_mpnn **out);
int os_mpnn_destroy(os_mpnn *net);
int os_mpnn_get_params(os_mpnn *net, float *params);
int os_mpnn_set_params(os_mpnn *net, const float *params);
/* marginals [count x n] of instances [first, first+count) of a dense batch */
int os_mpnn_predict(os_mpnn *net, os_dense *data, int first, int count, float *marginals);
/* mean binary cross-entropy over all spins and its gradient (packed order) */
int os_mpnn_loss_grad(os_mpnn *net, os_dense *data, int first, int count,
                      const uint8_t *labels, double *loss, float *grad);
/* one Adam step (beta1 0.9, beta2 0.999, eps 1e-8) on that loss */
int os_mpnn_train_step(os_mpnn *net, os_dense *data, int first, int count,
                       const uint8_t *labels, float lr, double *loss);

/* ---- batched sparse instances (CSR) ------------------------------------
 * Replaces, for many sparse instances at once, the loop a caller would write
 * around sa::anneal (include/simulated_annealing/annealing.hpp:55-140) and the
 * flatten of helpers::flatten_qubo (include/helpers/qubo_helpers.hpp:26-44):
 * the couplings stay sparse.  Instances are concatenated: instance i has n[i]
 * spins, n[i]+1 row_ptr entries (starting at 0, local offsets), row_ptr[n[i]]
 * col/val entries (local column indices, both (j,l) and (l,j) stored, i.e.
 * the symmetric off-diagonal part of the reference's flat matrix) and n[i]
 * diag entries.  Sweeps compute in fp32 (weights rounded from FP64); returned
 * energies are FP64 recomputations with the FP64 input. */
typedef struct os_csr_batch os_csr_batch;

typedef struct os_csr_params {
  uint32_t num_replicas;       /* replicas per instance */
  int32_t num_sweeps;          /* sweeps per replica */
  const float *tscale;         /* [num_sweeps] acceptance scale per sweep (1/beta) */
  uint64_t seed;               /* Philox key */
  uint32_t instance_base;      /* stream id of instance 0; instance i uses base + i (< 2^24) */
  const uint8_t *init_states;  /* optional start states, per instance [num_replicas x n[i]] */
  int32_t use_graph;           /* replay the run from a captured hipGraph */
  int32_t tune;                /* pick the launch shape by timing candidates on first use */
  int32_t site_order;          /* 0: every sweep visits sites 0..n-1 in order; 1: class by
                                  class of a greedy colouring (os_csr_batch_site_order), so a
                                  wave updates several independent sites of its replicas at
                                  once.  The update at position p of the visiting order draws
                                  word p&3 of counter (p>>2, replica, sweep, tag) either way */
  const float *init_marginals; /* optional warm start, per instance [n[i]] (instance-major):
                                  x_j = (u < p_j), u = (w >> 8) 2^-24 with w word j&3 of
                                  counter (j>>2, replica, 0, ((instance_base + i) << 8) | 0x12),
                                  the rule of os_sweep_params::init_marginals; sampled on the
                                  GPU.  Not together with init_states */
} os_csr_params;

/* Population annealing on a sparse batch (os_csr_batch_population): every
 * instance anneals num_replicas replicas through num_steps temperature
 * scales, sweeps_per_step sweeps each (sweep counters k * sweeps_per_step,
 * ...), with the thresholds and site order of os_csr_batch_anneal.  Each step
 * starts from the FP64 energy of its start states, as a run does.  After
 * every step but the last, the population is resampled systematically as in
 * os_population_params, on the running energies at the step's end and with
 * the uniform of Philox counter (k, 0, 0, ((instance_base + i) << 8) | 0x14).
 * The best state of replica index r is its lowest-energy state over the
 * steps (sweep ends and step starts; the earliest on ties).  No reference
 * counterpart: the reference recommends population annealing
 * (benchmarks/annealing/performance.md:52-59) but anneals only. */
typedef struct os_csr_population_params {
  uint32_t num_replicas;    /* population size R per instance */
  int32_t num_steps;        /* K >= 1 temperature steps */
  int32_t sweeps_per_step;  /* >= 1 */
  const float *tscale;      /* [num_steps] host array, > 0 */
  uint64_t seed;            /* Philox key */
  uint32_t instance_base;   /* stream id of instance 0 (< 2^24) */
  int32_t site_order;       /* as in os_csr_params */
  int32_t use_graph;        /* replay every step from the captured hipGraph */
  int32_t tune;             /* pick the launch shape by timing candidates on first use */
} os_csr_population_params;

int os_csr_batch_create(os_ctx *ctx, int32_t num_instances, const int32_t *n,
                        const int32_t *row_ptr, const int32_t *col, const double *val,
                        const double *diag, os_csr_batch **out);
int os_csr_batch_destroy(os_csr_batch *batch);
/* All outputs optional.  best_energies / run_best_energies: [num_instances x
 * num_replicas] FP64 energy of each best state / the sweep kernel's running
 * best; best_states / final_states: per instance [num_replicas x n[i]];
 * inst_best_energy / inst_best_replica: [num_instances] (first replica on
 * ties). */
int os_csr_batch_anneal(os_csr_batch *batch, const os_csr_params *params,
                        double *best_energies, double *run_best_energies, uint8_t *best_states,
                        uint8_t *final_states, double *inst_best_energy,
                        int32_t *inst_best_replica, os_sweep_stats *stats);
/* Marginals of the warm-start network on every spin of a sparse batch
 * (marginals: sum of n[i] floats, instance-major), the input of
 * os_csr_params::init_marginals.  The network reads each instance as it
 * reads the instance's dense matrix (features (Q_jj, sum_l Q_jl) / scale, 0, 1;
 * A = bf16(Q_jl / scale)), with a sparse aggregation; scale is the dense one
 * over scale_rows[i] rows (the instance's size in the dense layout the
 * network was trained on; NULL: n[i]).  The batch and the network must share
 * a context.  No reference counterpart. */
int os_mpnn_predict_csr(os_mpnn *net, os_csr_batch *batch, const int32_t *scale_rows,
                        float *marginals);
/* Replica exchange on a sparse batch (os_csr_batch_tempering): every
 * instance runs num_ladders ladders of num_temperatures slots; replica
 * r = l * num_temperatures + k starts in slot k of ladder l and sweeps at
 * the scale of its slot, with the thresholds and site order of
 * os_csr_batch_anneal.  Segments of exchange_interval sweeps each start from
 * the FP64 energy of their start states; after every segment but the last,
 * round t pairs slots (k, k+1), k = t mod 2, ..., of every ladder exactly as
 * os_dense_tempering does (u from Philox counter (k, l, t, ((instance_base +
 * i) << 8) | 0x13)), on the running energies at the segment's end.  The best
 * state of a replica is its lowest over the segments (sweep ends and segment
 * starts; the earliest on ties).  No reference counterpart: the reference
 * recommends replica exchange (benchmarks/annealing/performance.md:52-59)
 * but anneals only. */
typedef struct os_csr_tempering_params {
  uint32_t num_ladders;
  uint32_t num_temperatures; /* >= 2 slots per ladder */
  const float *tscale;       /* [num_temperatures] host array, > 0 */
  int32_t num_sweeps;        /* >= 1 */
  int32_t exchange_interval; /* sweeps between exchange rounds, >= 1 */
  uint64_t seed;             /* Philox key */
  uint32_t instance_base;    /* stream id of instance 0 (< 2^24) */
  int32_t site_order;        /* as in os_csr_params */
  int32_t use_graph;         /* replay every segment from the captured hipGraph */
  int32_t tune;              /* pick the launch shape by timing candidates on first use */
} os_csr_tempering_params;

/* Replica exchange (see os_csr_tempering_params).  Outputs optional, as in
 * os_csr_batch_anneal, indexed by replica; best_energies are the FP64
 * energies of the best states; final_slots: per instance [R] the slot of
 * every replica at the end. */
int os_csr_batch_tempering(os_csr_batch *batch, const os_csr_tempering_params *params,
                           double *best_energies, uint8_t *best_states, uint8_t *final_states,
                           uint32_t *final_slots, double *inst_best_energy,
                           int32_t *inst_best_replica, os_sweep_stats *stats);
/* Population annealing (see os_csr_population_params).  Outputs optional, as
 * in os_csr_batch_anneal; best_energies are the FP64 energies of the best
 * states; ancestors: per instance [num_steps - 1][num_replicas]. */
int os_csr_batch_population(os_csr_batch *batch, const os_csr_population_params *params,
                            double *best_energies, uint8_t *best_states, uint8_t *final_states,
                            uint32_t *ancestors, double *inst_best_energy,
                            int32_t *inst_best_replica, os_sweep_stats *stats);
/* The launch shape in use (0 before the first run); prefetch: 1 if the
 * kernel variant that requests the next site's row ahead runs (batches whose
 * rows have at most 8 couplings), 0 for the general one. */
int os_csr_batch_launch_shape(os_csr_batch *batch, int32_t *waves_per_group, int32_t *groups,
                              int32_t *prefetch);
/* The colour-class order's launch shape in use (0 before its first run).
 * A wave updates G sites of each of its 64/G replicas per step; every
 * instance has its own G (4, 8 or 16), chosen with the launch shape: larger
 * instances take more sites per step, so that their work items, which set
 * the length of a run, take fewer steps.  sites_per_step: the common G, or 0
 * when the instances differ.  OS_CSR_COLOUR_SITES (4, 8 or 16) in the
 * environment fixes one G for all. */
int os_csr_batch_colour_shape(os_csr_batch *batch, int32_t *waves_per_group, int32_t *groups,
                              int32_t *sites_per_step);
/* G of every instance ([num_instances], 0 before the first colour-class run). */
int os_csr_batch_colour_sites(os_csr_batch *batch, int32_t *sites_per_step);
/* The site order of site_order 1 for one instance (outputs optional):
 * order [n] lists the sites class by class, each class ascending;
 * class_start [num_classes + 1] bounds the classes.  Greedy colouring in
 * index order: site j takes the smallest colour held by none of its
 * neighbours l < j, so no two sites of a class are coupled. */
int os_csr_batch_site_order(os_csr_batch *batch, int32_t instance, int32_t *order,
                            int32_t *class_start, int32_t *num_classes);

/* ---- .qubo instances: text and binary ingest on the GPU ----
 *
 * An os_qubo holds a parsed model on the GPU: its distinct terms (i <= j, the
 * first value of a repeated key, in order of first appearance), n = largest
 * index + 1, and the header's max_nodes. */
typedef struct os_qubo os_qubo;

/* Parse `.qubo` text[0, len) -- replaces qubo::QUBOModel::load / parse_qubo
 * (include/model/qubo.hpp:392-430) and its QUBOBuilder checks (:282-379).
 * Accepts and rejects exactly what that grammar does; on rejection returns
 * OS_ERR_INVALID and os_last_error() is the reference's exception message
 * ("Parsing failed. Incorrect file format.", "An empty input, no coefficients
 * defined.", ...).  Numbers convert as strtod does (correctly rounded). */
int os_qubo_parse(os_ctx *ctx, const char *text, size_t len, os_qubo **out);
/* The same for text already in GPU memory (`text`: a device pointer on the
 * context's device, 16-byte aligned, len bytes readable).  No reference
 * counterpart: the reference parses from a host stream only. */
int os_qubo_parse_device(os_ctx *ctx, const void *text, size_t len, os_qubo **out);
/* Load the binary form written by os_qubo_write_binary (magic "OSQB",
 * version 1, little-endian: int32 n, int32 max_nodes, int64 num_linear,
 * int64 num_quadratic, then int32 i[T], int32 j[T], float64 v[T] with
 * T = num_linear + num_quadratic).  Rejects terms outside 0 <= i <= j < n,
 * repeated keys, counts that do not match, and n != largest index + 1.  No
 * reference counterpart (the reference reads text only). */
int os_qubo_read_binary(os_ctx *ctx, const void *data, size_t len, os_qubo **out);
int os_qubo_destroy(os_qubo *q);
/* Each output may be NULL.  n: QUBOModel::get_nodes (qubo.hpp:424 of the
 * builder: largest index + 1); num_linear / num_quadratic: the distinct
 * terms. */
int os_qubo_info(os_qubo *q, int32_t *n, int32_t *max_nodes, int64_t *num_linear,
                 int64_t *num_quadratic);
/* GPU time of the last text parse in ms: ms[0] line index, ms[1] line
 * parse, ms[2] dedupe and compaction. */
int os_qubo_parse_times(os_qubo *q, float *ms);
/* The terms (each output may be NULL; num_linear + num_quadratic entries). */
int os_qubo_terms(os_qubo *q, int32_t *i, int32_t *j, double *v);
/* helpers::flatten_qubo (include/helpers/qubo_helpers.hpp:26-44): the n x n
 * FP64 symmetric matrix into q_sym. */
int os_qubo_flatten(os_qubo *q, double *q_sym);
/* A dense sweep-mode problem built on the GPU from the terms (the flattened
 * matrix never passes through the host); n <= 16384. */
int os_qubo_dense(os_qubo *q, os_dense **out);
/* Bytes of the binary form (returns the size, not a status; -1 for NULL). */
int64_t os_qubo_binary_size(os_qubo *q);
int os_qubo_write_binary(os_qubo *q, void *buf, size_t len);

#ifdef __cplusplus
}
#endif
#endif /* ONESOLVER_H_ */
