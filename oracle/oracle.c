/*
 * oracle.c -- CPU restatement of the oneSolver simulated-annealing QUBO path.
 *
 * TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() (as the
 * checker) and bench.py's cpu_baseline leg.  The product never links it.
 * Compiled with -ffp-contract=off so that every floating-point expression is
 * evaluated exactly as written (explicit fma() where the specification
 * says fma).
 *
 * Reference files restated (paths relative to the oneSolver tree):
 *   include/simulated_annealing/annealing.hpp      sa::energy, sa::anneal
 *   include/simulated_annealing/device_rng.hpp     bit / bit_index / uniform
 *   include/helpers/qubo_helpers.hpp               flatten_qubo
 *   include/exhaustive/exhaustive.hpp              exhaustive::solve
 *   app/one-solver-anneal.cpp                      beta schedules
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------ */
/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as   */
/* easy as 1, 2, 3", SC'11).  Multipliers and Weyl key increments are the    */
/* published constants.                                                     */
/* ------------------------------------------------------------------------ */
#define PHILOX_M0 0xD2511F53u
#define PHILOX_M1 0xCD9E8D57u
#define PHILOX_W0 0x9E3779B9u
#define PHILOX_W1 0xBB67AE85u

void or_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
  uint32_t k0 = key[0], k1 = key[1];
  for (int round = 0; round < 10; ++round) {
    if (round > 0) {
      k0 += PHILOX_W0;
      k1 += PHILOX_W1;
    }
    uint64_t p0 = (uint64_t)PHILOX_M0 * c0;
    uint64_t p1 = (uint64_t)PHILOX_M1 * c2;
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    c0 = hi1 ^ c1 ^ k0;
    c1 = lo1;
    c2 = hi0 ^ c3 ^ k1;
    c3 = lo0;
  }
  out[0] = c0;
  out[1] = c1;
  out[2] = c2;
  out[3] = c3;
}

/* Stream tags (counter word 3, low byte); instance id in the upper 24 bits. */
enum {
  TAG_REF_INIT = 0x01,
  TAG_REF_STEP = 0x02,
  TAG_REF_STREAM = 0x03,
  TAG_SWEEP_INIT = 0x10,
  TAG_SWEEP_THR = 0x11,
  TAG_WARM = 0x12,
  TAG_EXCHANGE = 0x13,
  TAG_RESAMPLE = 0x14,
};

static void seed_key(uint64_t seed, uint32_t key[2]) {
  key[0] = (uint32_t)seed;
  key[1] = (uint32_t)(seed >> 32);
}

/* -ln(u) for u = (2*(w>>9)+1) * 2^-24, the open-interval 24-bit uniform of a
 * 32-bit word.  u = m * 2^e with m in [sqrt(1/2), sqrt(2)); ln(1+f) with
 * f = m-1 is f + f^2 * P(f), P a degree-6 least-squares fit in relative error
 * on that interval (max rel. error 2.2e-7 in fp32 evaluation).  Every step
 * is exact (frexp, *2, m-1) or a single correctly rounded fp32 operation, so
 * the value is reproducible bit for bit on any IEEE-754 implementation.      */
static const float LNP_C0 = -0x1.000080p-1f;
static const float LNP_C1 = 0x1.5557dep-2f;
static const float LNP_C2 = -0x1.ff5ac2p-3f;
static const float LNP_C3 = 0x1.97f830p-3f;
static const float LNP_C4 = -0x1.621a72p-3f;
static const float LNP_C5 = 0x1.4953eep-3f;
static const float LNP_C6 = -0x1.9190e0p-4f;
static const float LN2_F = 0x1.62e430p-1f;
static const float SQRT_HALF_F = 0x1.6a09e6p-1f;

float or_neg_ln_u(uint32_t w) {
  float u = (float)((w >> 9) * 2u + 1u) * 0x1p-24f;
  int e;
  float m = frexpf(u, &e);
  if (m < SQRT_HALF_F) {
    m = m + m;
    e = e - 1;
  }
  float f = m - 1.0f;
  float p = LNP_C6;
  p = fmaf(p, f, LNP_C5);
  p = fmaf(p, f, LNP_C4);
  p = fmaf(p, f, LNP_C3);
  p = fmaf(p, f, LNP_C2);
  p = fmaf(p, f, LNP_C1);
  p = fmaf(p, f, LNP_C0);
  float f2 = f * f;
  float l = fmaf(f2, p, f);
  l = fmaf((float)e, LN2_F, l);
  return -l;
}

void or_neg_ln_u_many(const uint32_t *w, int64_t count, float *out) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < count; ++i) out[i] = or_neg_ln_u(w[i]);
}

/* ------------------------------------------------------------------------ */
/* Model helpers                                                            */
/* ------------------------------------------------------------------------ */

/* include/helpers/qubo_helpers.hpp:26-44.  The model is given as the
 * contents of its two maps (unique keys).  Diagonal entries take
 * get_variable(i); every ordered pair (i,j), i!=j, adds get_connection(i,j)
 * to both [i + j*n] and [j + i*n], exactly as the reference loop does.       */
static double lookup_lin(int nlin, const int *idx, const double *val, int i) {
  for (int k = 0; k < nlin; ++k)
    if (idx[k] == i) return val[k];
  return 0.0;
}

void or_flatten(int n, int nlin, const int *lin_idx, const double *lin_val, int nquad,
                const int *q_i, const int *q_j, const double *q_val, double *out) {
  memset(out, 0, sizeof(double) * (size_t)n * (size_t)n);
  /* get_connection lookups: build a dense table of stored (i,j) keys. */
  double *conn = (double *)calloc((size_t)n * (size_t)n, sizeof(double));
  for (int k = 0; k < nquad; ++k) {
    if (q_i[k] >= 0 && q_i[k] < n && q_j[k] >= 0 && q_j[k] < n)
      conn[(size_t)q_i[k] * n + q_j[k]] = q_val[k];
  }
  for (int i = 0; i < n; ++i) {
    for (int j = 0; j < n; ++j) {
      if (i == j) {
        out[(size_t)i + (size_t)j * n] = lookup_lin(nlin, lin_idx, lin_val, i);
      } else {
        double c = conn[(size_t)i * n + j];
        out[(size_t)i + (size_t)j * n] += c;
        out[(size_t)j + (size_t)i * n] += c;
      }
    }
  }
  free(conn);
}

/* include/simulated_annealing/annealing.hpp:31-40: i ascending, j >= i
 * ascending, result += F[i*N+j] * x_i * x_j, in double.                     */
double or_energy(const double *flat, const uint8_t *state, int n) {
  double result = 0.0;
  for (int i = 0; i < n; i++) {
    for (int j = i; j < n; j++) {
      result += flat[(size_t)i * n + j] * (double)state[i] * (double)state[j];
    }
  }
  return result;
}

void or_energy_many(const double *flat, const uint8_t *states, int n, int64_t count,
                    double *out) {
#pragma omp parallel for schedule(static)
  for (int64_t r = 0; r < count; ++r) out[r] = or_energy(flat, states + (size_t)r * n, n);
}

/* app/one-solver-anneal.cpp:23-29.  beta_i = beta_min + beta_max*i/(n-1):
 * the last value is beta_min+beta_max, and n == 1 gives 0/0 = NaN.          */
void or_linear_schedule(double beta_min, double beta_max, uint32_t n, double *out) {
  for (uint32_t i = 0; i < n; i++) {
    out[i] = beta_min + beta_max * (double)i / (double)(uint32_t)(n - 1u);
  }
}

/* app/one-solver-anneal.cpp:31-39.  Cumulative product in double. */
void or_geometric_schedule(double beta_min, double beta_max, uint32_t n, double *out) {
  if (n == 0) return;
  out[0] = beta_min;
  double alpha = pow(beta_max / beta_min, 1.0 / (double)(uint32_t)(n - 1u));
  for (uint32_t i = 1; i < n; i++) out[i] = out[i - 1] * alpha;
}

/* ------------------------------------------------------------------------ */
/* Exhaustive search                                                        */
/* ------------------------------------------------------------------------ */

/* include/exhaustive/exhaustive.hpp:94-138 (per-state energy over the upper
 * triangle, strict '<', lowest index wins) and :158-166 (argmin).          */
int or_exhaustive_literal(int n, const double *flat, uint8_t *state, double *energy) {
  if (n < 1 || n > 24) return -1;
  uint64_t total = 1ull << n;
  double e_best = INFINITY;
  uint64_t s_best = 0;
  for (uint64_t s = 0; s < total; ++s) {
    double e = 0.0;
    for (int i = 0; i < n; ++i) {
      for (int j = i; j < n; ++j) {
        if (i == j) {
          if ((s >> i) & 1) e += flat[(size_t)i * n + i];
        } else {
          if (((s >> i) & 1) && ((s >> j) & 1)) e += flat[(size_t)i * n + j];
        }
      }
    }
    if (e < e_best) {
      e_best = e;
      s_best = s;
    }
  }
  for (int i = 0; i < n; ++i) state[i] = (uint8_t)((s_best >> i) & 1);
  *energy = e_best;
  return 0;
}

#define EXH_KEEP 16

typedef struct {
  double e;
  uint64_t s;
} cand_t;

static void cand_insert(cand_t *c, int *nc, double e, uint64_t s) {
  if (*nc == EXH_KEEP && e >= c[EXH_KEEP - 1].e) return;
  int k = (*nc < EXH_KEEP) ? (*nc)++ : EXH_KEEP - 1;
  while (k > 0 && c[k - 1].e > e) {
    c[k] = c[k - 1];
    --k;
  }
  c[k].e = e;
  c[k].s = s;
}

static double energy_bits(const double *flat, uint64_t s, int n) {
  double e = 0.0;
  for (int i = 0; i < n; ++i) {
    for (int j = i; j < n; ++j) {
      if (i == j) {
        if ((s >> i) & 1) e += flat[(size_t)i * n + i];
      } else if (((s >> i) & 1) && ((s >> j) & 1)) {
        e += flat[(size_t)i * n + j];
      }
    }
  }
  return e;
}

/* Gray-code enumeration with O(N) incremental energies.  The EXH_KEEP lowest
 * incremental energies of every chunk are re-ranked with the literal energy
 * (same summation order as the reference kernel) so that rounding of the
 * incremental sums cannot change the argmin or its tie order.               */
int or_exhaustive(int n, const double *flat, uint8_t *state, double *energy, int threads) {
  if (n < 1 || n > 40) return -1;
  uint64_t total = 1ull << n;
  int nchunks = 1;
#ifdef _OPENMP
  if (threads <= 0) threads = omp_get_max_threads();
  nchunks = threads * 8;
#else
  (void)threads;
#endif
  if ((uint64_t)nchunks > total) nchunks = (int)total;
  cand_t *all = (cand_t *)malloc(sizeof(cand_t) * EXH_KEEP * (size_t)nchunks);
  int *counts = (int *)calloc((size_t)nchunks, sizeof(int));
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads > 0 ? threads : 1)
  for (int c = 0; c < nchunks; ++c) {
    uint64_t rem = total % (uint64_t)nchunks, uc = (uint64_t)c;
    uint64_t g0 = total / (uint64_t)nchunks * uc + (uc < rem ? uc : rem);
    uint64_t g1 = g0 + total / (uint64_t)nchunks + (uc < rem ? 1u : 0u);
    cand_t *cand = all + (size_t)c * EXH_KEEP;
    int nc = 0;
    double *h = (double *)malloc(sizeof(double) * n);
    uint64_t s = g0 ^ (g0 >> 1);
    for (int j = 0; j < n; ++j) {
      double acc = flat[(size_t)j * n + j];
      for (int l = 0; l < n; ++l)
        if (l != j && ((s >> l) & 1)) acc += flat[(size_t)j * n + l];
      h[j] = acc;
    }
    double e = energy_bits(flat, s, n);
    cand_insert(cand, &nc, e, s);
    for (uint64_t g = g0 + 1; g < g1; ++g) {
      int b = __builtin_ctzll(g);
      int xb = (int)((s >> b) & 1);
      e += xb ? -h[b] : h[b];
      double d = xb ? -1.0 : 1.0;
      s ^= 1ull << b;
      for (int l = 0; l < n; ++l)
        if (l != b) h[l] += flat[(size_t)l * n + b] * d;
      cand_insert(cand, &nc, e, s);
    }
    counts[c] = nc;
    free(h);
  }
  double e_best = INFINITY;
  uint64_t s_best = 0;
  double e_min_inc = INFINITY;
  for (int c = 0; c < nchunks; ++c)
    if (counts[c] > 0 && all[(size_t)c * EXH_KEEP].e < e_min_inc) e_min_inc = all[(size_t)c * EXH_KEEP].e;
  double tol = 1e-9 * (fabs(e_min_inc) + 1.0);
  for (int c = 0; c < nchunks; ++c) {
    for (int k = 0; k < counts[c]; ++k) {
      cand_t cd = all[(size_t)c * EXH_KEEP + k];
      if (cd.e > e_min_inc + tol) continue;
      double ex = energy_bits(flat, cd.s, n);
      if (ex < e_best || (ex == e_best && cd.s < s_best)) {
        e_best = ex;
        s_best = cd.s;
      }
    }
  }
  for (int i = 0; i < n; ++i) state[i] = (uint8_t)((s_best >> i) & 1);
  *energy = e_best;
  free(all);
  free(counts);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Reference-mode annealer                                                  */
/* ------------------------------------------------------------------------ */

/* device_rng.hpp:46 bit(): spin j of replica i from counter
 * (j>>7, i, 0, TAG_REF_INIT), word (j>>5)&3, bit j&31.                     */
static void ref_init_bits(int n, uint32_t replica, const uint32_t key[2], uint8_t *x) {
  uint32_t out[4];
  for (int j = 0; j < n; ++j) {
    if ((j & 127) == 0) {
      uint32_t ctr[4] = {(uint32_t)(j >> 7), replica, 0u, TAG_REF_INIT};
      or_philox4x32_10(ctr, key, out);
    }
    x[j] = (uint8_t)((out[(j >> 5) & 3] >> (j & 31)) & 1u);
  }
}

/* device_rng.hpp:54-56 bit_index() and :62 uniform(): attempt t of replica
 * i uses counter (t_lo, i, t_hi, TAG_REF_STEP); site = floor(w0*N/2^32),
 * u = w1 * 2^-32 in [0,1).                                                  */
static void ref_attempt(uint64_t t, uint32_t replica, const uint32_t key[2], int n, int *site,
                        double *u) {
  uint32_t ctr[4] = {(uint32_t)t, replica, (uint32_t)(t >> 32), TAG_REF_STEP};
  uint32_t out[4];
  or_philox4x32_10(ctr, key, out);
  *site = (int)(((uint64_t)out[0] * (uint64_t)(uint32_t)n) >> 32);
  *u = (double)out[1] * 0x1p-32;
}

/* annealing.hpp:85-126, one work-item per replica, literal O(N^2) energy
 * recompute after every proposed flip (:104).                               */
void or_anneal_literal(int n, const double *flat, const double *beta, int num_iter,
                       int sweeps_per_beta, uint32_t num_tries, uint64_t seed,
                       uint32_t first_replica, uint8_t *best_states, double *best_energies,
                       int threads, uint8_t *final_states) {
  uint32_t key[2];
  seed_key(seed, key);
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads > 0 ? threads : 1)
  for (int64_t ii = 0; ii < (int64_t)num_tries; ++ii) {
    uint32_t i = first_replica + (uint32_t)ii;
    uint8_t *x = (uint8_t *)malloc((size_t)n);
    uint8_t *best = best_states + (size_t)ii * n;
    ref_init_bits(n, i, key, x);
    memcpy(best, x, (size_t)n);
    double best_energy = or_energy(flat, x, n);
    double current_energy = best_energy;
    uint64_t t = 0;
    for (int iter = 0; iter < num_iter; iter++) {
      double b = beta[iter];
      for (int sweep = 0; sweep < sweeps_per_beta; sweep++) {
        int s;
        double u;
        ref_attempt(t++, i, key, n, &s, &u);
        x[s] = (uint8_t)(1 - x[s]);
        double new_energy = or_energy(flat, x, n);
        if ((new_energy < current_energy) || (exp((current_energy - new_energy) / b) > u)) {
          current_energy = new_energy;
        } else {
          x[s] = (uint8_t)(1 - x[s]);
        }
        if (current_energy < best_energy) {
          best_energy = current_energy;
          memcpy(best, x, (size_t)n);
        }
      }
    }
    best_energies[ii] = best_energy;
    if (final_states) memcpy(final_states + (size_t)ii * n, x, (size_t)n);
    free(x);
  }
}

/* Same Markov chain as or_anneal_literal with local fields
 * h_j = F_jj + sum_{l!=j} F_jl x_l, dE = (1-2x_s) h_s, E tracked by
 * E += dE; the acceptance test is annealing.hpp:106-107 with
 * E_cur - E_new = -dE.                                                      */
void or_anneal_fields(int n, const double *flat, const double *beta, int num_iter,
                      int sweeps_per_beta, uint32_t num_tries, uint64_t seed,
                      uint32_t first_replica, uint8_t *best_states, double *best_energies,
                      int threads, uint8_t *final_states) {
  uint32_t key[2];
  seed_key(seed, key);
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads > 0 ? threads : 1)
  for (int64_t ii = 0; ii < (int64_t)num_tries; ++ii) {
    uint32_t i = first_replica + (uint32_t)ii;
    uint8_t *x = (uint8_t *)malloc((size_t)n);
    double *h = (double *)malloc(sizeof(double) * (size_t)n);
    uint8_t *best = best_states + (size_t)ii * n;
    ref_init_bits(n, i, key, x);
    memcpy(best, x, (size_t)n);
    for (int j = 0; j < n; ++j) {
      double acc = flat[(size_t)j * n + j];
      for (int l = 0; l < n; ++l)
        if (l != j && x[l]) acc += flat[(size_t)j * n + l];
      h[j] = acc;
    }
    double best_energy = or_energy(flat, x, n);
    double current_energy = best_energy;
    uint64_t t = 0;
    for (int iter = 0; iter < num_iter; iter++) {
      double b = beta[iter];
      for (int sweep = 0; sweep < sweeps_per_beta; sweep++) {
        int s;
        double u;
        ref_attempt(t++, i, key, n, &s, &u);
        double de = x[s] ? -h[s] : h[s];
        if ((de < 0.0) || (exp(-de / b) > u)) {
          double d = x[s] ? -1.0 : 1.0;
          x[s] = (uint8_t)(1 - x[s]);
          for (int l = 0; l < n; ++l)
            if (l != s) h[l] += flat[(size_t)l * n + s] * d;
          current_energy += de;
        }
        if (current_energy < best_energy) {
          best_energy = current_energy;
          memcpy(best, x, (size_t)n);
        }
      }
    }
    best_energies[ii] = best_energy;
    if (final_states) memcpy(final_states + (size_t)ii * n, x, (size_t)n);
    free(h);
    free(x);
  }
}

/* One shared 32-bit stream read at per-replica offsets, as the reference's
 * engines are built (annealing.hpp:87: philox4x32x10 engine(1234, i), i.e.
 * replica i starts i outputs into one seed's stream, device_rng.hpp:26-30).
 * Word p of the stream is word p%4 of Philox counter (p/4 lo, p/4 hi, 0,
 * TAG_REF_STREAM).  Each scalar draw consumes the next word:
 *   bit()       = w >> 31                          (device_rng.hpp:46)
 *   bit_index() = floor(w * N / 2^32)              (:55)
 *   uniform()   = w * 2^-32, from uniform_words words, the last one used
 *                 (:62; 1 or 2: how many outputs oneMKL's uniform<double>
 *                 takes is not known here)
 * and uniform() is drawn only on the branch that is not downhill (the
 * short-circuit || of annealing.hpp:106-107).  The word-to-value transforms
 * are this restatement's choice; the per-replica law equals
 * or_anneal_fields', and the replicas share words as the reference's do.
 * The num_tries replicas form independent runs of block_tries replicas each
 * (0: one run), run b on the stream of seed + b.                            */
static uint32_t stream_word(uint64_t p, const uint32_t key[2]) {
  uint32_t ctr[4] = {(uint32_t)(p >> 2), (uint32_t)(p >> 34), 0u, TAG_REF_STREAM};
  uint32_t out[4];
  or_philox4x32_10(ctr, key, out);
  return out[p & 3];
}

void or_anneal_stream(int n, const double *flat, const double *beta, int num_iter,
                      int sweeps_per_beta, uint32_t num_tries, uint32_t block_tries,
                      uint64_t seed, int uniform_words, uint8_t *best_states,
                      double *best_energies, int threads, uint8_t *final_states) {
  if (block_tries == 0) block_tries = num_tries;
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads > 0 ? threads : 1)
  for (int64_t ii = 0; ii < (int64_t)num_tries; ++ii) {
    /* blocks of block_tries replicas, block b on the stream of seed + b */
    uint32_t key[2];
    seed_key(seed + (uint64_t)ii / block_tries, key);
    uint64_t pos = (uint64_t)ii % block_tries;
    uint8_t *x = (uint8_t *)malloc((size_t)n);
    double *h = (double *)malloc(sizeof(double) * (size_t)n);
    uint8_t *best = best_states + (size_t)ii * n;
    for (int j = 0; j < n; ++j) x[j] = (uint8_t)(stream_word(pos++, key) >> 31);
    memcpy(best, x, (size_t)n);
    for (int j = 0; j < n; ++j) {
      double acc = flat[(size_t)j * n + j];
      for (int l = 0; l < n; ++l)
        if (l != j && x[l]) acc += flat[(size_t)j * n + l];
      h[j] = acc;
    }
    double best_energy = or_energy(flat, x, n);
    double current_energy = best_energy;
    for (int iter = 0; iter < num_iter; iter++) {
      double b = beta[iter];
      for (int sweep = 0; sweep < sweeps_per_beta; sweep++) {
        const int s = (int)(((uint64_t)stream_word(pos++, key) * (uint64_t)(uint32_t)n) >> 32);
        double de = x[s] ? -h[s] : h[s];
        int accept = de < 0.0;
        if (!accept) {
          double u = 0.0;
          for (int k = 0; k < uniform_words; ++k) u = (double)stream_word(pos++, key) * 0x1p-32;
          accept = exp(-de / b) > u;
        }
        if (accept) {
          double d = x[s] ? -1.0 : 1.0;
          x[s] = (uint8_t)(1 - x[s]);
          for (int l = 0; l < n; ++l)
            if (l != s) h[l] += flat[(size_t)l * n + s] * d;
          current_energy += de;
        }
        if (current_energy < best_energy) {
          best_energy = current_energy;
          memcpy(best, x, (size_t)n);
        }
      }
    }
    best_energies[ii] = best_energy;
    if (final_states) memcpy(final_states + (size_t)ii * n, x, (size_t)n);
    free(h);
    free(x);
  }
}

/* ------------------------------------------------------------------------ */
/* Sweep-mode annealers                                                     */
/* ------------------------------------------------------------------------ */

static void sweep_init_bits(int n, uint32_t replica, uint32_t instance, const uint32_t key[2],
                            uint8_t *x) {
  uint32_t out[4];
  for (int j = 0; j < n; ++j) {
    if ((j & 127) == 0) {
      uint32_t ctr[4] = {(uint32_t)(j >> 7), replica, 0u, (instance << 8) | TAG_SWEEP_INIT};
      or_philox4x32_10(ctr, key, out);
    }
    x[j] = (uint8_t)((out[(j >> 5) & 3] >> (j & 31)) & 1u);
  }
}

void or_sweep_init(int n, uint32_t num_replicas, uint64_t seed, uint32_t instance,
                   uint8_t *states) {
  uint32_t key[2];
  seed_key(seed, key);
  for (uint32_t r = 0; r < num_replicas; ++r)
    sweep_init_bits(n, r, instance, key, states + (size_t)r * n);
}

void or_warm_init(int n, const float *p, uint32_t num_replicas, uint64_t seed,
                  uint32_t instance, uint8_t *states) {
  uint32_t key[2];
  seed_key(seed, key);
  uint32_t out[4];
  for (uint32_t r = 0; r < num_replicas; ++r) {
    for (int j = 0; j < n; ++j) {
      if ((j & 3) == 0) {
        uint32_t ctr[4] = {(uint32_t)(j >> 2), r, 0u, (instance << 8) | TAG_WARM};
        or_philox4x32_10(ctr, key, out);
      }
      float u = (float)(out[j & 3] >> 8) * 0x1p-24f;
      states[(size_t)r * n + j] = (uint8_t)(u < p[j] ? 1 : 0);
    }
  }
}

/* Threshold words of sweep t for replica r: counter (j>>2, r, t, tag). */
static void sweep_thr_words(uint32_t j4, uint32_t r, uint32_t t, uint32_t instance,
                            const uint32_t key[2], uint32_t out[4]) {
  uint32_t ctr[4] = {j4, r, t, (instance << 8) | TAG_SWEEP_THR};
  or_philox4x32_10(ctr, key, out);
}

void or_sweep_dense(int n, const double *flat, const float *tscale, int num_sweeps,
                    uint32_t num_replicas, uint64_t seed, uint32_t instance,
                    const uint8_t *init, uint8_t *final_states, uint8_t *best_states,
                    double *best_energies, int threads) {
  uint32_t key[2];
  seed_key(seed, key);
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads > 0 ? threads : 1)
  for (int64_t rr = 0; rr < (int64_t)num_replicas; ++rr) {
    uint32_t r = (uint32_t)rr;
    uint8_t *x = (uint8_t *)malloc((size_t)n);
    double *h = (double *)malloc(sizeof(double) * (size_t)n);
    uint8_t *best = best_states + (size_t)r * n;
    if (init)
      memcpy(x, init + (size_t)r * n, (size_t)n);
    else
      sweep_init_bits(n, r, instance, key, x);
    for (int j = 0; j < n; ++j) {
      double acc = flat[(size_t)j * n + j];
      for (int l = 0; l < n; ++l)
        if (l != j && x[l]) acc += flat[(size_t)j * n + l];
      h[j] = acc;
    }
    double e = or_energy(flat, x, n);
    double e_best = e;
    memcpy(best, x, (size_t)n);
    uint32_t w[4];
    for (int t = 0; t < num_sweeps; ++t) {
      float ts = tscale[t];
      for (int j = 0; j < n; ++j) {
        if ((j & 3) == 0) sweep_thr_words((uint32_t)(j >> 2), r, (uint32_t)t, instance, key, w);
        float thr = or_neg_ln_u(w[j & 3]) * ts;
        double de = x[j] ? -h[j] : h[j];
        if (de < (double)thr) {
          double d = x[j] ? -1.0 : 1.0;
          x[j] = (uint8_t)(1 - x[j]);
          for (int l = 0; l < n; ++l)
            if (l != j) h[l] += flat[(size_t)l * n + j] * d;
          e += de;
        }
      }
      if (e < e_best) {
        e_best = e;
        memcpy(best, x, (size_t)n);
      }
    }
    if (final_states) memcpy(final_states + (size_t)r * n, x, (size_t)n);
    best_energies[r] = e_best;
    free(h);
    free(x);
  }
}

/* Systematic resampling of one population of R replicas: weights
 * w_i = exp(-dbeta (E_i - E_min)), cumulative sums W_i in replica order;
 * new replica j descends from the smallest i with W_i > (j + u) / R * W_{R-1}
 * (u in [0, 1)).  Shared by the oracle's population annealer and its tests. */
void or_resample(const double *E, uint32_t R, double dbeta, double u, uint32_t *anc) {
  double emin = E[0];
  for (uint32_t i = 1; i < R; ++i) emin = E[i] < emin ? E[i] : emin;
  double *W = (double *)malloc(sizeof(double) * R);
  double acc = 0.0;
  for (uint32_t i = 0; i < R; ++i) {
    acc += exp(-dbeta * (E[i] - emin));
    W[i] = acc;
  }
  uint32_t i = 0;
  for (uint32_t j = 0; j < R; ++j) {
    const double target = ((double)j + u) / (double)R * W[R - 1];
    while (i + 1 < R && W[i] <= target) ++i;
    anc[j] = i;
  }
  free(W);
}

/* The uniform of resampling round k: words 0 and 1 of Philox counter
 * (k, 0, 0, (instance << 8) | TAG_RESAMPLE), 53 bits.                       */
double or_resample_uniform(uint64_t seed, uint32_t instance, uint32_t k) {
  uint32_t key[2];
  seed_key(seed, key);
  uint32_t ctr[4] = {k, 0u, 0u, (instance << 8) | TAG_RESAMPLE};
  uint32_t o[4];
  or_philox4x32_10(ctr, key, o);
  return (double)(((uint64_t)(o[0] >> 5) << 26) | (uint64_t)(o[1] >> 6)) * 0x1p-53;
}

/* Population annealing over sweep-mode replicas (os_dense_population, no
 * reference counterpart; the reference recommends it,
 * benchmarks/annealing/performance.md:52-59).  R replicas start from the
 * sweep-mode random bits.  Step k = 0..K-1 runs theta sweeps at threshold
 * scale tscale[k] (sweep counters k*theta ..); after every step but the last
 * the population is resampled (or_resample) with dbeta = 1/tscale[k+1] -
 * 1/tscale[k], the running energies and u = or_resample_uniform(k): replica
 * j then continues from the state, fields and running energy of replica
 * anc[j].  The best state of each replica index is tracked at sweep ends and
 * never moved.  ancestors: optional [R x (K-1)] ancestor lists per round.   */
void or_population_dense(int n, const double *flat, const float *tscale, int K, int theta,
                         uint32_t R, uint64_t seed, uint32_t instance, uint8_t *final_states,
                         uint8_t *best_states, double *best_energies, uint32_t *ancestors,
                         int threads) {
  uint32_t key[2];
  seed_key(seed, key);
  uint8_t *xs = (uint8_t *)malloc((size_t)R * n);
  double *hs = (double *)malloc(sizeof(double) * (size_t)R * n);
  double *es = (double *)malloc(sizeof(double) * R);
  uint8_t *xs2 = (uint8_t *)malloc((size_t)R * n);
  double *hs2 = (double *)malloc(sizeof(double) * (size_t)R * n);
  double *es2 = (double *)malloc(sizeof(double) * R);
  uint32_t *anc = (uint32_t *)malloc(sizeof(uint32_t) * R);
  for (uint32_t r = 0; r < R; ++r) {
    uint8_t *x = xs + (size_t)r * n;
    double *h = hs + (size_t)r * n;
    sweep_init_bits(n, r, instance, key, x);
    for (int j = 0; j < n; ++j) {
      double acc = flat[(size_t)j * n + j];
      for (int l = 0; l < n; ++l)
        if (l != j && x[l]) acc += flat[(size_t)j * n + l];
      h[j] = acc;
    }
    es[r] = or_energy(flat, x, n);
    best_energies[r] = es[r];
    memcpy(best_states + (size_t)r * n, x, (size_t)n);
  }
  for (int k = 0; k < K; ++k) {
    const float ts = tscale[k];
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads > 0 ? threads : 1)
    for (int64_t rr = 0; rr < (int64_t)R; ++rr) {
      const uint32_t r = (uint32_t)rr;
      uint8_t *x = xs + (size_t)r * n;
      double *h = hs + (size_t)r * n;
      uint32_t w[4];
      for (int t = k * theta; t < (k + 1) * theta; ++t) {
        for (int j = 0; j < n; ++j) {
          if ((j & 3) == 0) sweep_thr_words((uint32_t)(j >> 2), r, (uint32_t)t, instance, key, w);
          float thr = or_neg_ln_u(w[j & 3]) * ts;
          double de = x[j] ? -h[j] : h[j];
          if (de < (double)thr) {
            double d = x[j] ? -1.0 : 1.0;
            x[j] = (uint8_t)(1 - x[j]);
            for (int l = 0; l < n; ++l)
              if (l != j) h[l] += flat[(size_t)l * n + j] * d;
            es[r] += de;
          }
        }
        if (es[r] < best_energies[r]) {
          best_energies[r] = es[r];
          memcpy(best_states + (size_t)r * n, x, (size_t)n);
        }
      }
    }
    if (k + 1 >= K) break;
    const double dbeta = 1.0 / (double)tscale[k + 1] - 1.0 / (double)tscale[k];
    or_resample(es, R, dbeta, or_resample_uniform(seed, instance, (uint32_t)k), anc);
    if (ancestors) memcpy(ancestors + (size_t)k * R, anc, sizeof(uint32_t) * R);
    for (uint32_t j = 0; j < R; ++j) {
      memcpy(xs2 + (size_t)j * n, xs + (size_t)anc[j] * n, (size_t)n);
      memcpy(hs2 + (size_t)j * n, hs + (size_t)anc[j] * n, sizeof(double) * (size_t)n);
      es2[j] = es[anc[j]];
    }
    uint8_t *tx = xs; xs = xs2; xs2 = tx;
    double *th = hs; hs = hs2; hs2 = th;
    double *te = es; es = es2; es2 = te;
  }
  if (final_states) memcpy(final_states, xs, (size_t)R * n);
  free(xs); free(hs); free(es); free(xs2); free(hs2); free(es2); free(anc);
}

/* Replica exchange over sweep-mode replicas (os_dense_tempering, no reference
 * counterpart).  Replica r = l*K + k of ladder l starts in slot k; it sweeps
 * with the thresholds of or_sweep_dense at scale tscale[slot].  After every
 * `interval` sweeps (but not after the last), round t pairs slots (k, k+1),
 * k = t%2, t%2+2, ...: with E the running energies, the replicas swap slots
 * iff d = (1/ts_k - 1/ts_{k+1}) (E_k - E_{k+1}) >= 0 or u < exp(d), u = (w>>8)
 * 2^-24 from word 0 of Philox counter (k, l, t, (instance << 8) | 0x13).      */
void or_tempering_dense(int n, const double *flat, const float *tscale, int K,
                        uint32_t ladders, int num_sweeps, int interval, uint64_t seed,
                        uint32_t instance, uint8_t *final_states, uint8_t *best_states,
                        double *best_energies, uint32_t *final_slots, int threads) {
  uint32_t key[2];
  seed_key(seed, key);
  const uint32_t R = ladders * (uint32_t)K;
  uint8_t *xs = (uint8_t *)malloc((size_t)R * n);
  double *hs = (double *)malloc(sizeof(double) * (size_t)R * n);
  double *es = (double *)malloc(sizeof(double) * R);
  uint32_t *slot = (uint32_t *)malloc(sizeof(uint32_t) * R);
  uint32_t *at = (uint32_t *)malloc(sizeof(uint32_t) * R);
  for (uint32_t r = 0; r < R; ++r) {
    uint8_t *x = xs + (size_t)r * n;
    double *h = hs + (size_t)r * n;
    sweep_init_bits(n, r, instance, key, x);
    for (int j = 0; j < n; ++j) {
      double acc = flat[(size_t)j * n + j];
      for (int l = 0; l < n; ++l)
        if (l != j && x[l]) acc += flat[(size_t)j * n + l];
      h[j] = acc;
    }
    es[r] = or_energy(flat, x, n);
    best_energies[r] = es[r];
    memcpy(best_states + (size_t)r * n, x, (size_t)n);
    slot[r] = r % (uint32_t)K;
    at[r] = r;
  }
  uint32_t round = 0;
  for (int s0 = 0; s0 < num_sweeps; s0 += interval) {
    const int ns = num_sweeps - s0 < interval ? num_sweeps - s0 : interval;
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads > 0 ? threads : 1)
    for (int64_t rr = 0; rr < (int64_t)R; ++rr) {
      const uint32_t r = (uint32_t)rr;
      uint8_t *x = xs + (size_t)r * n;
      double *h = hs + (size_t)r * n;
      const float ts = tscale[slot[r]];
      uint32_t w[4];
      for (int t = s0; t < s0 + ns; ++t) {
        for (int j = 0; j < n; ++j) {
          if ((j & 3) == 0) sweep_thr_words((uint32_t)(j >> 2), r, (uint32_t)t, instance, key, w);
          float thr = or_neg_ln_u(w[j & 3]) * ts;
          double de = x[j] ? -h[j] : h[j];
          if (de < (double)thr) {
            double d = x[j] ? -1.0 : 1.0;
            x[j] = (uint8_t)(1 - x[j]);
            for (int l = 0; l < n; ++l)
              if (l != j) h[l] += flat[(size_t)l * n + j] * d;
            es[r] += de;
          }
        }
        if (es[r] < best_energies[r]) {
          best_energies[r] = es[r];
          memcpy(best_states + (size_t)r * n, x, (size_t)n);
        }
      }
    }
    if (s0 + ns >= num_sweeps) break;
    for (uint32_t l = 0; l < ladders; ++l) {
      uint32_t *ra = at + (size_t)l * K;
      for (int k = (int)(round & 1u); k + 1 < K; k += 2) {
        const uint32_t a = ra[k], b = ra[k + 1];
        const double d = (1.0 / (double)tscale[k] - 1.0 / (double)tscale[k + 1]) * (es[a] - es[b]);
        uint32_t ctr[4] = {(uint32_t)k, l, round, (instance << 8) | TAG_EXCHANGE};
        uint32_t o[4];
        or_philox4x32_10(ctr, key, o);
        const double u = (double)(o[0] >> 8) * 0x1p-24;
        if (d >= 0.0 || u < exp(d)) {
          ra[k] = b;
          ra[k + 1] = a;
          slot[a] = (uint32_t)(k + 1);
          slot[b] = (uint32_t)k;
        }
      }
    }
    ++round;
  }
  if (final_states) memcpy(final_states, xs, (size_t)R * n);
  if (final_slots) memcpy(final_slots, slot, sizeof(uint32_t) * R);
  free(xs);
  free(hs);
  free(es);
  free(slot);
  free(at);
}

/* Energy of a CSR instance in the fixed order used by the GPU kernels:
 * E = sum_j x_j * (diag_j + 0.5 * sum_{k in row j} val_k x_col_k), double. */
static double csr_energy(int n, const int32_t *row_ptr, const int32_t *col, const float *val,
                         const float *diag, const uint8_t *x) {
  double e = 0.0;
  for (int j = 0; j < n; ++j) {
    if (!x[j]) continue;
    double off = 0.0;
    for (int k = row_ptr[j]; k < row_ptr[j + 1]; ++k)
      if (x[col[k]]) off += (double)val[k];
    e += (double)diag[j] + 0.5 * off;
  }
  return e;
}

void or_sweep_csr(int n, const int32_t *row_ptr, const int32_t *col, const float *val,
                  const float *diag, const float *tscale, int num_sweeps,
                  uint32_t num_replicas, uint64_t seed, uint32_t instance,
                  const uint8_t *init, uint8_t *final_states, uint8_t *best_states,
                  double *best_energies, int threads) {
  or_sweep_csr_order(n, row_ptr, col, val, diag, tscale, num_sweeps, num_replicas, seed,
                     instance, NULL, init, final_states, best_states, best_energies, threads);
}

/* Colour classes of a symmetric CSR graph: greedy in index order, site j
 * taking the smallest colour no neighbour l < j holds.  order[] lists the
 * sites by (colour, index); class c is order[cstart[c] .. cstart[c+1]).
 * Returns the number of colours (cstart needs max degree + 2 entries).      */
int or_csr_colour_order(int n, const int32_t *row_ptr, const int32_t *col, int32_t *order,
                        int32_t *cstart) {
  int32_t *colour = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
  int maxdeg = 0;
  for (int j = 0; j < n; ++j)
    if (row_ptr[j + 1] - row_ptr[j] > maxdeg) maxdeg = row_ptr[j + 1] - row_ptr[j];
  uint8_t *used = (uint8_t *)calloc((size_t)maxdeg + 2, 1);
  int ncol = 0;
  for (int j = 0; j < n; ++j) {
    for (int k = row_ptr[j]; k < row_ptr[j + 1]; ++k)
      if (col[k] < j) used[colour[col[k]]] = 1;
    int c = 0;
    while (used[c]) ++c;
    colour[j] = c;
    if (c + 1 > ncol) ncol = c + 1;
    for (int k = row_ptr[j]; k < row_ptr[j + 1]; ++k)
      if (col[k] < j) used[colour[col[k]]] = 0;
  }
  for (int c = 0; c <= ncol; ++c) cstart[c] = 0;
  for (int j = 0; j < n; ++j) ++cstart[colour[j] + 1];
  for (int c = 0; c < ncol; ++c) cstart[c + 1] += cstart[c];
  int32_t *fill = (int32_t *)malloc(sizeof(int32_t) * (size_t)(ncol + 1));
  memcpy(fill, cstart, sizeof(int32_t) * (size_t)(ncol + 1));
  for (int j = 0; j < n; ++j) order[fill[colour[j]]++] = j;
  free(fill);
  free(used);
  free(colour);
  return ncol;
}

/* The sparse sweep with the sites of each sweep visited in order[0..n-1]
 * (NULL: 0..n-1).  The threshold of the update at position pos of a sweep
 * is word pos&3 of Philox counter (pos>>2, r, t): for the order 0..n-1 the
 * stream of or_sweep_csr.                                                  */
/* Sweeps t0 .. t0+num_sweeps-1 of replica r from state x (updated in place)
 * at scale tscale[t] (tscale NULL: ts_fixed for every sweep): the energy
 * starts as csr_energy(x); best/e_best take the start state and every sweep
 * end below the best so far.  Returns the running energy.                  */
static double csr_chain(int n, const int32_t *row_ptr, const int32_t *col, const float *val,
                        const float *diag, const float *tscale, float ts_fixed, int t0,
                        int num_sweeps, uint32_t r, uint32_t instance, const uint32_t key[2],
                        const int32_t *order, uint8_t *x, uint8_t *best, double *e_best) {
  double e = csr_energy(n, row_ptr, col, val, diag, x);
  double eb = e;
  memcpy(best, x, (size_t)n);
  uint32_t w[4];
  for (int t = t0; t < t0 + num_sweeps; ++t) {
    float ts = tscale ? tscale[t] : ts_fixed;
    for (int pos = 0; pos < n; ++pos) {
      const int j = order ? order[pos] : pos;
      if ((pos & 3) == 0) sweep_thr_words((uint32_t)(pos >> 2), r, (uint32_t)t, instance, key, w);
      float thr = or_neg_ln_u(w[pos & 3]) * ts;
      float h = diag[j];
      for (int k = row_ptr[j]; k < row_ptr[j + 1]; ++k)
        if (x[col[k]]) h = h + val[k];
      float de = x[j] ? -h : h;
      if (de < thr) {
        x[j] = (uint8_t)(1 - x[j]);
        e += (double)de;
      }
    }
    if (e < eb) {
      eb = e;
      memcpy(best, x, (size_t)n);
    }
  }
  *e_best = eb;
  return e;
}

void or_sweep_csr_order(int n, const int32_t *row_ptr, const int32_t *col, const float *val,
                        const float *diag, const float *tscale, int num_sweeps,
                        uint32_t num_replicas, uint64_t seed, uint32_t instance,
                        const int32_t *order, const uint8_t *init, uint8_t *final_states,
                        uint8_t *best_states, double *best_energies, int threads) {
  uint32_t key[2];
  seed_key(seed, key);
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads > 0 ? threads : 1)
  for (int64_t rr = 0; rr < (int64_t)num_replicas; ++rr) {
    uint32_t r = (uint32_t)rr;
    uint8_t *x = (uint8_t *)malloc((size_t)n);
    if (init)
      memcpy(x, init + (size_t)r * n, (size_t)n);
    else
      sweep_init_bits(n, r, instance, key, x);
    csr_chain(n, row_ptr, col, val, diag, tscale, 0.0f, 0, num_sweeps, r, instance, key, order,
              x, best_states + (size_t)r * n, best_energies + r);
    if (final_states) memcpy(final_states + (size_t)r * n, x, (size_t)n);
    free(x);
  }
}

/* Population annealing over sparse sweep-mode replicas (os_csr_batch_population,
 * no reference counterpart).  Step k = 0..K-1 runs csr_chain over sweeps
 * k*theta .. at tscale[k] from the current states (step 0: the sweep-mode
 * random bits); the best of replica index r is the step's best where it is
 * below the best so far (step 0: taken as is).  After every step but the
 * last, or_resample with dbeta = 1/tscale[k+1] - 1/tscale[k], the running
 * energies at the step's end and u = or_resample_uniform(k) picks the state
 * replica j continues from.  ancestors: optional [(K-1) x R].               */
void or_population_csr(int n, const int32_t *row_ptr, const int32_t *col, const float *val,
                       const float *diag, const float *tscale, int K, int theta, uint32_t R,
                       uint64_t seed, uint32_t instance, const int32_t *order,
                       uint8_t *final_states, uint8_t *best_states, double *best_energies,
                       uint32_t *ancestors, int threads) {
  uint32_t key[2];
  seed_key(seed, key);
  float *ts = (float *)malloc(sizeof(float) * (size_t)K * theta);
  for (int k = 0; k < K; ++k)
    for (int t = 0; t < theta; ++t) ts[(size_t)k * theta + t] = tscale[k];
  uint8_t *xs = (uint8_t *)malloc((size_t)R * n);
  uint8_t *xs2 = (uint8_t *)malloc((size_t)R * n);
  uint8_t *sb = (uint8_t *)malloc((size_t)R * n);
  double *es = (double *)malloc(sizeof(double) * R);
  double *seb = (double *)malloc(sizeof(double) * R);
  uint32_t *anc = (uint32_t *)malloc(sizeof(uint32_t) * R);
  for (uint32_t r = 0; r < R; ++r) sweep_init_bits(n, r, instance, key, xs + (size_t)r * n);
  for (int k = 0; k < K; ++k) {
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads > 0 ? threads : 1)
    for (int64_t rr = 0; rr < (int64_t)R; ++rr) {
      const uint32_t r = (uint32_t)rr;
      es[r] = csr_chain(n, row_ptr, col, val, diag, ts, 0.0f, k * theta, theta, r, instance, key,
                        order, xs + (size_t)r * n, sb + (size_t)r * n, seb + r);
    }
    for (uint32_t r = 0; r < R; ++r)
      if (k == 0 || seb[r] < best_energies[r]) {
        best_energies[r] = seb[r];
        memcpy(best_states + (size_t)r * n, sb + (size_t)r * n, (size_t)n);
      }
    if (k + 1 >= K) break;
    const double dbeta = 1.0 / (double)tscale[k + 1] - 1.0 / (double)tscale[k];
    or_resample(es, R, dbeta, or_resample_uniform(seed, instance, (uint32_t)k), anc);
    if (ancestors) memcpy(ancestors + (size_t)k * R, anc, sizeof(uint32_t) * R);
    for (uint32_t j = 0; j < R; ++j)
      memcpy(xs2 + (size_t)j * n, xs + (size_t)anc[j] * n, (size_t)n);
    uint8_t *tx = xs; xs = xs2; xs2 = tx;
  }
  if (final_states) memcpy(final_states, xs, (size_t)R * n);
  free(ts); free(xs); free(xs2); free(sb); free(es); free(seb); free(anc);
}

/* Replica exchange over sparse sweep-mode replicas (os_csr_batch_tempering,
 * no reference counterpart).  Replica r = l*K + k of ladder l starts in
 * slot k.  Segments of `interval` sweeps (the last may be shorter) run
 * csr_chain at tscale[slot[r]] from the current states (segment 0: the
 * sweep-mode random bits); the best of replica r is the segment's best
 * where it is below the best so far (segment 0: taken as is).  After every
 * segment but the last, round t pairs slots (k, k+1), k = t%2, t%2+2, ...
 * of every ladder, with the running energies at the segment's end, exactly
 * as or_tempering_dense does.  num_sweeps >= 1.                            */
void or_tempering_csr(int n, const int32_t *row_ptr, const int32_t *col, const float *val,
                      const float *diag, const float *tscale, int K, uint32_t ladders,
                      int num_sweeps, int interval, uint64_t seed, uint32_t instance,
                      const int32_t *order, uint8_t *final_states, uint8_t *best_states,
                      double *best_energies, uint32_t *final_slots, int threads) {
  uint32_t key[2];
  seed_key(seed, key);
  const uint32_t R = ladders * (uint32_t)K;
  uint8_t *xs = (uint8_t *)malloc((size_t)R * n);
  uint8_t *sb = (uint8_t *)malloc((size_t)R * n);
  double *es = (double *)malloc(sizeof(double) * R);
  double *seb = (double *)malloc(sizeof(double) * R);
  uint32_t *slot = (uint32_t *)malloc(sizeof(uint32_t) * R);
  uint32_t *at = (uint32_t *)malloc(sizeof(uint32_t) * R);
  for (uint32_t r = 0; r < R; ++r) {
    sweep_init_bits(n, r, instance, key, xs + (size_t)r * n);
    slot[r] = r % (uint32_t)K;
    at[r] = r;
  }
  uint32_t round = 0;
  for (int s0 = 0; s0 < num_sweeps; s0 += interval) {
    const int ns = num_sweeps - s0 < interval ? num_sweeps - s0 : interval;
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads > 0 ? threads : 1)
    for (int64_t rr = 0; rr < (int64_t)R; ++rr) {
      const uint32_t r = (uint32_t)rr;
      es[r] = csr_chain(n, row_ptr, col, val, diag, NULL, tscale[slot[r]], s0, ns, r, instance,
                        key, order, xs + (size_t)r * n, sb + (size_t)r * n, seb + r);
    }
    for (uint32_t r = 0; r < R; ++r)
      if (s0 == 0 || seb[r] < best_energies[r]) {
        best_energies[r] = seb[r];
        memcpy(best_states + (size_t)r * n, sb + (size_t)r * n, (size_t)n);
      }
    if (s0 + ns >= num_sweeps) break;
    for (uint32_t l = 0; l < ladders; ++l) {
      uint32_t *ra = at + (size_t)l * K;
      for (int k = (int)(round & 1u); k + 1 < K; k += 2) {
        const uint32_t a = ra[k], b = ra[k + 1];
        const double d = (1.0 / (double)tscale[k] - 1.0 / (double)tscale[k + 1]) * (es[a] - es[b]);
        uint32_t ctr[4] = {(uint32_t)k, l, round, (instance << 8) | TAG_EXCHANGE};
        uint32_t o[4];
        or_philox4x32_10(ctr, key, o);
        const double u = (double)(o[0] >> 8) * 0x1p-24;
        if (d >= 0.0 || u < exp(d)) {
          ra[k] = b;
          ra[k + 1] = a;
          slot[a] = (uint32_t)(k + 1);
          slot[b] = (uint32_t)k;
        }
      }
    }
    ++round;
  }
  if (final_states) memcpy(final_states, xs, (size_t)R * n);
  if (final_slots) memcpy(final_slots, slot, sizeof(uint32_t) * R);
  free(xs); free(sb); free(es); free(seb); free(slot); free(at);
}

/* ------------------------------------------------------------------------
 * .qubo reader: parse_qubo + QUBOBuilder, include/model/qubo.hpp:282-417.
 *
 *   file        := comment* header? (comment | coefficient)*   (:406)
 *   comment     := 'c' print* (eol | eoi)                       (:399)
 *   header      := "p qubo" uint uint uint uint (eol | eoi)     (:400-404)
 *   coefficient := index index real (eol | eoi)                 (:396-398)
 *   index       := lexeme[uint ' ']                             (:395)
 * with blanks (space, tab) skipped before every token (qi::blank, :407) and
 * the whole input consumed (:410).  add_element runs as soon as the three
 * numbers are read, before the end of the line is checked (:397), and
 * throws for i > j; numbers are converted by strtod (correctly rounded).
 * ------------------------------------------------------------------------ */
typedef struct {
  const char *p, *end;
  /* builder state, qubo.hpp:284-289 */
  int num_quadratic, num_linear, max_nodes;
  int have_used, max_used;
  int64_t cap, count, nlin, nquad;
  int32_t *ti, *tj;
  double *tv;
  uint64_t *hkeys; /* open addressing over (i, j), first value kept */
  uint64_t hmask;
  int lower;       /* add_element threw */
} QP;

static int qp_blank(int c) { return c == ' ' || c == '\t'; }
static int qp_at(const QP *s, const char *q) { return q < s->end ? (unsigned char)*q : -1; }
static void qp_skip(QP *s) {
  while (s->p < s->end && qp_blank((unsigned char)*s->p)) ++s->p;
}
static int qp_eol_or_eoi(QP *s) {
  qp_skip(s);
  int c = qp_at(s, s->p);
  if (c < 0) return 1;
  if (c == '\n') { ++s->p; return 1; }
  if (c == '\r') {
    ++s->p;
    if (qp_at(s, s->p) == '\n') ++s->p;
    return 1;
  }
  return 0;
}
/* qi::uint_: decimal digits, no sign, fails beyond 2^32 - 1 */
static int qp_uint(QP *s, const char **q, unsigned *out) {
  const char *t = *q;
  int c = qp_at(s, t);
  if (c < '0' || c > '9') return 0;
  uint64_t v = 0;
  while ((c = qp_at(s, t)) >= '0' && c <= '9') {
    v = v * 10 + (uint64_t)(c - '0');
    if (v > 0xFFFFFFFFull) return 0;
    ++t;
  }
  *out = (unsigned)v;
  *q = t;
  return 1;
}
static int qp_lower(int c) { return c >= 'A' && c <= 'Z' ? c + 32 : c; }
static int qp_word_ci(QP *s, const char **q, const char *w) {
  const char *t = *q;
  for (; *w; ++w, ++t)
    if (qp_lower(qp_at(s, t)) != *w) return 0;
  *q = t;
  return 1;
}
/* qi::double_: sign? (nan ("(" ... ")")? | inf (inity)? | digits ["." digits]
 * | "." digits) (e sign? digits)? -- the exponent only when it has digits. */
static int qp_real(QP *s, double *out) {
  const char *t = s->p, *start = s->p;
  int neg = 0, c = qp_at(s, t);
  if (c == '+' || c == '-') { neg = c == '-'; ++t; c = qp_at(s, t); }
  if (c == 'n' || c == 'N') {
    if (!qp_word_ci(s, &t, "nan")) return 0;
    if (qp_at(s, t) == '(') {
      const char *u = t + 1;
      while (u < s->end && *u != ')') ++u;
      if (u < s->end) t = u + 1;
    }
    *out = neg ? -(double)NAN : (double)NAN;
    s->p = t;
    return 1;
  }
  if (c == 'i' || c == 'I') {
    if (!qp_word_ci(s, &t, "inf")) return 0;
    qp_word_ci(s, &t, "inity");
    *out = neg ? -(double)INFINITY : (double)INFINITY;
    s->p = t;
    return 1;
  }
  int digits = 0;
  while ((c = qp_at(s, t)) >= '0' && c <= '9') { ++t; digits = 1; }
  if (qp_at(s, t) == '.') {
    const char *u = t + 1;
    int frac = 0;
    while ((c = qp_at(s, u)) >= '0' && c <= '9') { ++u; frac = 1; }
    if (digits || frac) { t = u; digits = 1; }
  }
  if (!digits) return 0;
  c = qp_at(s, t);
  if (c == 'e' || c == 'E') {
    const char *u = t + 1;
    c = qp_at(s, u);
    if (c == '+' || c == '-') ++u;
    int ed = 0;
    while ((c = qp_at(s, u)) >= '0' && c <= '9') { ++u; ed = 1; }
    if (ed) t = u;
  }
  size_t L = (size_t)(t - start);
  char buf[64], *tok = L < sizeof buf ? buf : (char *)malloc(L + 1);
  memcpy(tok, start, L);
  tok[L] = 0;
  *out = strtod(tok, NULL);
  if (tok != buf) free(tok);
  s->p = t;
  return 1;
}
static int qp_comment(QP *s) {
  const char *save = s->p;
  qp_skip(s);
  if (qp_at(s, s->p) != 'c') { s->p = save; return 0; }
  ++s->p;
  for (;;) {
    qp_skip(s);
    int c = qp_at(s, s->p);
    if (c >= 0x20 && c < 0x7F) ++s->p; else break;   /* qi::print */
  }
  if (!qp_eol_or_eoi(s)) { s->p = save; return 0; }
  return 1;
}
static void qp_header(QP *s) {
  const char *save = s->p;
  qp_skip(s);
  const char *q = s->p;
  for (const char *w = "p qubo"; *w; ++w, ++q)
    if (qp_at(s, q) != *w) { s->p = save; return; }
  s->p = q;
  for (int k = 0; k < 4; ++k) {
    unsigned v;
    qp_skip(s);
    if (!qp_uint(s, &s->p, &v)) { s->p = save; return; }
    if (k == 1) s->max_nodes = (int)v;      /* the actions run as each number is read */
    if (k == 2) s->num_linear = (int)v;
    if (k == 3) s->num_quadratic = (int)v;
  }
  if (!qp_eol_or_eoi(s)) s->p = save;
}
static uint64_t qp_hash(uint64_t k) { return (k * 0x9E3779B97F4A7C15ull) >> 17; }
/* QUBOBuilder::add_element, qubo.hpp:304-318 */
static int qp_add(QP *s, int i, int j, double v) {
  if (i > j) { s->lower = 1; return 0; }
  uint64_t key = ((uint64_t)(uint32_t)i << 32) | (uint32_t)j;
  uint64_t h = qp_hash(key) & s->hmask;
  for (;;) {
    if (s->hkeys[h] == key + 1) break;   /* repeated key: the first value stays */
    if (s->hkeys[h] == 0) {
      if (s->count >= s->cap) return 0;
      s->hkeys[h] = key + 1;
      s->ti[s->count] = i;
      s->tj[s->count] = j;
      s->tv[s->count] = v;
      ++s->count;
      if (i == j) ++s->nlin; else ++s->nquad;
      break;
    }
    h = (h + 1) & s->hmask;
  }
  if (!s->have_used || j > s->max_used) s->max_used = j;   /* std::set<int>::rbegin */
  if (i > s->max_used) s->max_used = i;
  s->have_used = 1;
  return 1;
}
static int qp_index(QP *s, unsigned *out) {
  qp_skip(s);
  const char *q = s->p;
  if (!qp_uint(s, &q, out)) return 0;
  if (qp_at(s, q) != ' ') return 0;
  s->p = q + 1;
  return 1;
}
static int qp_coefficient(QP *s) {
  const char *save = s->p;
  unsigned i, j;
  double v;
  if (!qp_index(s, &i) || !qp_index(s, &j)) { s->p = save; return 0; }
  qp_skip(s);
  if (!qp_real(s, &v)) { s->p = save; return 0; }
  if (!qp_add(s, (int)i, (int)j, v)) return 0;
  if (!qp_eol_or_eoi(s)) { s->p = save; return 0; }
  return 1;
}

int or_qubo_parse(const char *text, int64_t len, int64_t cap, int32_t *ti, int32_t *tj,
                  double *tv, int64_t *nt, int32_t *n, int64_t *num_linear,
                  int64_t *num_quadratic) {
  QP s;
  memset(&s, 0, sizeof s);
  s.p = text;
  s.end = text + len;
  s.num_quadratic = s.num_linear = s.max_nodes = -1;
  s.cap = cap;
  s.ti = ti;
  s.tj = tj;
  s.tv = tv;
  uint64_t hc = 16;
  while (hc < 2 * (uint64_t)cap + 16) hc <<= 1;
  s.hkeys = (uint64_t *)calloc(hc, sizeof(uint64_t));
  s.hmask = hc - 1;
  while (qp_comment(&s)) {
  }
  qp_header(&s);
  while (qp_comment(&s) || qp_coefficient(&s)) {
  }
  free(s.hkeys);
  if (s.lower) return OR_QP_LOWER;
  qp_skip(&s);
  if (s.p != s.end) return OR_QP_SYNTAX;
  /* QUBOBuilder::build_qubo, qubo.hpp:353-377 */
  if (s.nlin == 0 && s.nquad == 0) return OR_QP_EMPTY;
  if (s.num_quadratic == -1 || s.num_linear == -1 || s.max_nodes == -1) return OR_QP_NO_HEADER;
  if (s.num_linear > s.max_nodes) return OR_QP_LINEAR_GT_NODES;
  /* size_t != int: the int converts to size_t (sign-extended) */
  if ((uint64_t)s.nquad != (uint64_t)(int64_t)s.num_quadratic) return OR_QP_QUAD_COUNT;
  if ((uint64_t)s.nlin != (uint64_t)(int64_t)s.num_linear) return OR_QP_LINEAR_COUNT;
  *nt = s.count;
  *n = s.max_used + 1;
  *num_linear = s.nlin;
  *num_quadratic = s.nquad;
  return OR_QP_OK;
}
