/*
 * oracle.h -- CPU restatement of the oneSolver simulated-annealing QUBO path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (onesolver_amd/, the
 * one-solver-anneal CLI, the C-ABI library) links, loads or calls this code.
 * It is used by tests/, by __graft_entry__.smoke() as the checker, and by
 * bench.py's cpu_baseline leg as the timed CPU baseline.
 *
 * Every function cites the reference file:line it restates.  Where the
 * reference's behaviour depends on an unavailable third-party library
 * (oneMKL device RNG, see include/simulated_annealing/device_rng.hpp:26-30),
 * the oracle uses the counter-based Philox4x32-10 stream layout that the
 * product uses, so that trajectories can be compared bit for bit; the law of
 * each replica's Markov chain is the reference's.
 *
 * Parity pinning (how each function is checked before it is trusted):
 *   or_philox4x32_10   Random123 / SC'11 known-answer vectors, and rocRAND's
 *                      philox4x32_10 header compiled on the host.
 *   or_neg_ln_u        libm log over the whole 2^23-point input grid.
 *   or_energy          reference KATs tests/exhaustive_test.cpp:12-142 via
 *                      or_exhaustive_*; published CSV rows of
 *                      benchmarks/annealing/results/128/001.csv.
 *   or_flatten         tests/qubo_helpers_test.cpp:14-27 (25 expected values).
 *   or_*_schedule      closed forms (no reference test exists for them).
 *   or_exhaustive_*    reference KATs (-12, -1.2, -22, -14, -32) and brute
 *                      force cross-check literal vs Gray code.
 *   or_anneal_literal  invariants (returned energy == or_energy of returned
 *                      state; >= exhaustive ground energy), bit-identical to
 *                      or_anneal_fields on integer instances, hit-rate
 *                      against the exhaustive ground state.
 *   or_sweep_dense /   exact greedy limit (beta -> inf reduces to
 *   or_sweep_csr       deterministic descent, checked against a brute-force
 *                      local-minimum test), detailed-balance check of the
 *                      visited-state distribution on N=4 at fixed beta
 *                      against the Boltzmann law.
 *   or_mpnn_*          finite differences of the FP64 loss.
 *   or_qubo_parse      the accept/reject strings and expected models of the
 *                      reference's tests/io_test.cpp:12-124 and of the product
 *                      reader's tests; the example files' models (flattened
 *                      and compared with the reference's
 *                      tests/qubo_helpers_test.cpp:14-27 expectation and the
 *                      exhaustive KAT energies); numbers against Python's
 *                      float(), an independent correctly rounded conversion.
 */
#ifndef ONESOLVER_ORACLE_H_
#define ONESOLVER_ORACLE_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- RNG specification shared with the product (see oracle.c) ---------- */
void or_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
float or_neg_ln_u(uint32_t w);
void or_neg_ln_u_many(const uint32_t *w, int64_t count, float *out);

/* ---- model helpers ------------------------------------------------------- */
/* include/helpers/qubo_helpers.hpp:26-44 */
void or_flatten(int n, int nlin, const int *lin_idx, const double *lin_val,
                int nquad, const int *q_i, const int *q_j, const double *q_val,
                double *out);
/* include/simulated_annealing/annealing.hpp:31-40 */
double or_energy(const double *flat, const uint8_t *state, int n);
void or_energy_many(const double *flat, const uint8_t *states, int n,
                    int64_t count, double *out);
/* app/one-solver-anneal.cpp:23-29 and :31-39 */
void or_linear_schedule(double beta_min, double beta_max, uint32_t n, double *out);
void or_geometric_schedule(double beta_min, double beta_max, uint32_t n, double *out);

/* ---- exhaustive search (include/exhaustive/exhaustive.hpp:29-167) -------- */
/* Literal O(N^2)/state restatement, N <= 24. Returns 0 on success. */
int or_exhaustive_literal(int n, const double *flat, uint8_t *state, double *energy);
/* Gray-code O(N)/state with exact re-ranking of near-ties, N <= 40. */
int or_exhaustive(int n, const double *flat, uint8_t *state, double *energy, int threads);

/* ---- reference-mode annealer (include/simulated_annealing/annealing.hpp:55-140) */
/* Per replica i in [first_replica, first_replica+num_tries): init bits, then
 * num_iter*sweeps_per_beta random-site attempts with acceptance
 * E_new < E_cur || exp((E_cur-E_new)/beta) > u.
 * literal: O(N^2) energy recompute per attempt (annealing.hpp:104).
 * fields : same chain with O(N) incremental local fields.                 */
void or_anneal_literal(int n, const double *flat, const double *beta, int num_iter,
                       int sweeps_per_beta, uint32_t num_tries, uint64_t seed,
                       uint32_t first_replica, uint8_t *best_states,
                       double *best_energies, int threads, uint8_t *final_states);
void or_anneal_fields(int n, const double *flat, const double *beta, int num_iter,
                      int sweeps_per_beta, uint32_t num_tries, uint64_t seed,
                      uint32_t first_replica, uint8_t *best_states,
                      double *best_energies, int threads, uint8_t *final_states);

/* The reference's replicas read one stream at offsets 0, 1, 2, ...
 * (annealing.hpp:87); this restates that sharing, with the chain of
 * or_anneal_fields.  uniform_words: stream words per uniform() draw.
 * Replicas come in runs of block_tries (0: one run of num_tries); run b
 * uses the stream of seed + b.                                            */
void or_anneal_stream(int n, const double *flat, const double *beta, int num_iter,
                      int sweeps_per_beta, uint32_t num_tries, uint32_t block_tries,
                      uint64_t seed, int uniform_words, uint8_t *best_states,
                      double *best_energies, int threads, uint8_t *final_states);

/* ---- sweep-mode annealers (configs with "sweeps": sequential site order) -- */
/* tscale[t]: threshold scale of sweep t (1/beta_t for inverse-temperature
 * schedules, beta_t when beta is used as a temperature as in
 * annealing.hpp:107).  A spin flips iff dE < neg_ln_u(w) * tscale[t].
 * init: NULL -> Philox random bits; else init[r*n + j] gives the start state.
 * Best state is tracked at sweep boundaries (and at the start).            */
void or_sweep_dense(int n, const double *flat, const float *tscale, int num_sweeps,
                    uint32_t num_replicas, uint64_t seed, uint32_t instance,
                    const uint8_t *init, uint8_t *final_states, uint8_t *best_states,
                    double *best_energies, int threads);
/* Replica exchange (os_dense_tempering): num_ladders * K replicas, replica
 * l*K + k starting in slot k, sweeping at tscale[slot]; exchange rounds of
 * adjacent slots after every `interval` sweeps (not after the last).       */
/* Population annealing (os_dense_population): K steps of theta sweeps at
 * tscale[k]; systematic resampling between steps with weights
 * exp(-(1/tscale[k+1] - 1/tscale[k]) (E_i - E_min)) and the uniform of
 * or_resample_uniform; ancestors optional [(K-1) x R].                     */
void or_resample(const double *E, uint32_t R, double dbeta, double u, uint32_t *anc);
double or_resample_uniform(uint64_t seed, uint32_t instance, uint32_t k);
void or_population_dense(int n, const double *flat, const float *tscale, int K, int theta,
                         uint32_t R, uint64_t seed, uint32_t instance, uint8_t *final_states,
                         uint8_t *best_states, double *best_energies, uint32_t *ancestors,
                         int threads);
void or_tempering_dense(int n, const double *flat, const float *tscale, int K,
                        uint32_t ladders, int num_sweeps, int interval, uint64_t seed,
                        uint32_t instance, uint8_t *final_states, uint8_t *best_states,
                        double *best_energies, uint32_t *final_slots, int threads);
/* CSR over the off-diagonal couplings (symmetric), fp32 values, fp32 diag. */
void or_sweep_csr(int n, const int32_t *row_ptr, const int32_t *col, const float *val,
                  const float *diag, const float *tscale, int num_sweeps,
                  uint32_t num_replicas, uint64_t seed, uint32_t instance,
                  const uint8_t *init, uint8_t *final_states, uint8_t *best_states,
                  double *best_energies, int threads);
/* or_sweep_csr with the sites of every sweep visited in order[0..n-1]
 * (NULL: 0..n-1); the update at position pos draws word pos&3 of counter
 * (pos>>2, r, t), or_sweep_csr's stream for the order 0..n-1.              */
void or_sweep_csr_order(int n, const int32_t *row_ptr, const int32_t *col, const float *val,
                        const float *diag, const float *tscale, int num_sweeps,
                        uint32_t num_replicas, uint64_t seed, uint32_t instance,
                        const int32_t *order, const uint8_t *init, uint8_t *final_states,
                        uint8_t *best_states, double *best_energies, int threads);
/* Population annealing over sparse sweep-mode replicas (os_csr_batch_population):
 * steps of theta sweeps of or_sweep_csr_order's chains at tscale[k], each
 * from the FP64-recomputed energy of its start states; systematic
 * resampling (or_resample) between steps; best per replica index over the
 * steps.  ancestors: optional [(K-1) x R].                                 */
void or_population_csr(int n, const int32_t *row_ptr, const int32_t *col, const float *val,
                       const float *diag, const float *tscale, int K, int theta, uint32_t R,
                       uint64_t seed, uint32_t instance, const int32_t *order,
                       uint8_t *final_states, uint8_t *best_states, double *best_energies,
                       uint32_t *ancestors, int threads);
/* Replica exchange over sparse sweep-mode replicas (os_csr_batch_tempering):
 * segments of `interval` sweeps of or_sweep_csr_order's chains at the scale
 * of each replica's slot, each from the FP64-recomputed energy of its start
 * states; or_tempering_dense's exchange rounds between segments; best per
 * replica over the segments.                                                */
void or_tempering_csr(int n, const int32_t *row_ptr, const int32_t *col, const float *val,
                      const float *diag, const float *tscale, int K, uint32_t ladders,
                      int num_sweeps, int interval, uint64_t seed, uint32_t instance,
                      const int32_t *order, uint8_t *final_states, uint8_t *best_states,
                      double *best_energies, uint32_t *final_slots, int threads);
/* Greedy colour classes of a symmetric CSR graph (index order, smallest
 * colour free of lower-index neighbours); order[] by (colour, index),
 * class c = order[cstart[c] .. cstart[c+1]); returns the colour count.     */
int or_csr_colour_order(int n, const int32_t *row_ptr, const int32_t *col, int32_t *order,
                        int32_t *cstart);
/* Warm-start initial state from marginals p[j]: x_j = (u < p_j). */
void or_warm_init(int n, const float *p, uint32_t num_replicas, uint64_t seed,
                  uint32_t instance, uint8_t *states);
/* Random initial bits of sweep mode (shared by dense and CSR kernels). */
void or_sweep_init(int n, uint32_t num_replicas, uint64_t seed, uint32_t instance,
                   uint8_t *states);

/* Outcomes of or_qubo_parse: 0 or one of the reference's std::invalid_argument
 * throws (include/model/qubo.hpp:313-314, :355-373, :413-415). */
enum {
  OR_QP_OK = 0,
  OR_QP_SYNTAX = 1,           /* "Parsing failed. Incorrect file format." */
  OR_QP_LOWER = 2,            /* "... coefficient from lower triangle ..." */
  OR_QP_EMPTY = 3,            /* "An empty input, no coefficients defined." */
  OR_QP_NO_HEADER = 4,        /* "No header line or header line misformatted." */
  OR_QP_LINEAR_GT_NODES = 5,  /* "Number of linear terms is greater than num nodes." */
  OR_QP_QUAD_COUNT = 6,       /* "Number of quadratic terms is not equal ..." */
  OR_QP_LINEAR_COUNT = 7      /* "Number of linear terms is not equal ..." */
};

/* parse_qubo + QUBOBuilder (include/model/qubo.hpp:282-417) over text[0, len):
 * the distinct terms in order of first appearance (ti <= tj; the first value
 * of a repeated key is kept), at most `cap` of them; *n = max index + 1.
 * Returns an OR_QP_* code; the outputs are valid for OR_QP_OK only. */
int or_qubo_parse(const char *text, int64_t len, int64_t cap, int32_t *ti, int32_t *tj,
                  double *tv, int64_t *nt, int32_t *n, int64_t *num_linear,
                  int64_t *num_quadratic);

#ifdef __cplusplus
}
#endif
#endif
