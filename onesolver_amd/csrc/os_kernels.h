// os_kernels.h -- internal launchers of the HIP kernels (not part of the C ABI).
#pragma once

#include "os_common.h"

namespace os {

// energy.hip
// bit-packed states: row pitch W words (W >= ceil(n/32)); bits >= n are zero
void launch_pack_bytes(hipStream_t st, const uint8_t *bytes, int n, int count, int W,
                       uint32_t *bits);
void launch_unpack_bits(hipStream_t st, const uint32_t *bits, int n, int count, int W,
                        uint8_t *bytes);
size_t energy_scratch_bytes(int n, int count);
// out[r] = FP64 energy of bit-packed state r.  With `fields` non-null, also
// refreshes fields[r*ldf + j] (fp32) and raises *drift_max (fp32 bit pattern).
// num_instances matrices (stride n*n) with `count` consecutive states each.
void launch_energy_fp64(hipStream_t st, int n, int W, const double *q, int count,
                        const uint32_t *bits, double *out, double *scratch, float *fields, int ldf,
                        unsigned int *drift_max, int num_instances = 1);

// capi_noalloc.hip: looks up this device's address of the fixed scratch
// array (once per device; the device must be current)
void noalloc_setup(int device);

// anneal_ref.hip
int ref_lds_limit();
size_t ref_lds_bytes(int n);
void launch_ref_anneal(hipStream_t st, int n, const double *q, const double *beta, int num_iter,
                       int spb, uint32_t num_tries, uint64_t seed, double *gfields,
                       uint32_t *best_bits, double *best_run);

// exhaustive.hip
int exhaustive_max_n();
// entries of the per-block partial buffers, and of the per-thread minima,
// that exhaustive_solve needs for n
size_t exhaustive_partials(int n);
size_t exhaustive_thread_slots(int n);
// minimum over all 2^n states of the energy of the upper triangle of q_dev
// (n x n FP64, device), lowest index on ties; `finite`: every used
// coefficient is finite.  Synchronises the stream.
void exhaustive_solve(hipStream_t st, int n, const double *q_dev, bool finite, double *tmin,
                      double *part_e, uint64_t *part_x, size_t part_cap, double *best_e,
                      uint64_t *best_x);

// anneal_dense.hip
int dense_max_n();
int dense_padded_n(int n);
// num_instances instances of n spins (qpack / qdb stacked per instance), R_pad
// replicas each (instance i at rows [i*R_pad, (i+1)*R_pad), stream id
// instance + i)
// tsr: optional per-row temperature scales (replica exchange), else tscale
// per sweep
void launch_dense_sweeps(hipStream_t st, int n, int np, int num_instances, const uint4 *qpack,
                         const float *qdb, const float *tscale, const float *tsr, int sweep0,
                         int nsweeps, uint64_t seed, uint32_t instance, int R_pad, float *fields,
                         uint32_t *xbits, uint32_t *bestbits, double *energy, double *ebest,
                         int *status);
// population annealing: rows r < R of every instance (Rp rows each) take
// the state words, fields and energy of row anc[instance][r]; rows R..Rp-1
// keep theirs (out-of-place: x2, f2, e2)
void launch_dense_gather(hipStream_t st, int num_instances, int Rp, int R, int npw, int np,
                         const uint32_t *anc, const uint32_t *x, const float *f, const double *e,
                         uint32_t *x2, float *f2, double *e2);
// start states of num_instances instances (R rows each, stream ids instance + i;
// warm start: marginals p[i][n])
void launch_sweep_init(hipStream_t st, int n, int npw, int R, uint64_t seed, uint32_t instance,
                       uint32_t *xbits, int num_instances = 1);
void launch_warm_init(hipStream_t st, int n, int npw, int R, uint64_t seed, uint32_t instance,
                      const float *p, uint32_t *xbits, int num_instances = 1);

// anneal_stream.hip (dense n > dense_max_n(): fields in HBM, blocks of 64)
int stream_block();
void launch_pack_q64(hipStream_t st, int I, int n, int np, const double *q, uint16_t *qpack,
                     float *qdb);
// Blocks g of a segment are numbered from 0 (block g covers spins 64*(g mod
// NB).. of sweep g / NB); the flip masks of block g live in dring[g % ring]
// (dring: ring x I*Rp uint4).
// pass launch: Fout = Fin + the flips of blocks g0 .. g0+cnt-1, the operands
// of `chunk` (1, 2 or 4) blocks staged per round trip, `grid` persistent
// work-groups
void launch_stream_pass(hipStream_t st, int I, int np, int Rp, int g0, int cnt, int ring,
                        int chunk, int grid, const uint4 *qpack, const float *Fin, float *Fout,
                        const uint4 *dring);
// decide launch: block g, from the fields Fin (after blocks .. gb-1) plus the
// flips of blocks gb .. g-1
// variant: 10 * (pending blocks loaded per round trip) + (1 if the next
// chunk's loads are issued before the current chunk is applied), or 60 + c
// for the shared-operand form (waves 0 and 1 add the pending flips to both
// replica halves of a column half, c = 0, 1, 2: 2, 1, 3 blocks per round
// trip; waves 2 and 3 make the thresholds); 61 default
void launch_stream_decide(hipStream_t st, int I, int n, int np, int Rp, int g, int gb, int ring,
                          int sweep, float ts, uint64_t seed, uint32_t inst_base,
                          bool last_of_sweep, const uint4 *qpack, const float *qdb,
                          const float *Fin, uint32_t *xbits, uint4 *dring, double *energy,
                          double *ebest, uint32_t *bestbits, int variant);

// gen.hip
// q[s] (FP64 n x n, s < I) = SK instance inst_base + s of key seed
void launch_gen_sk(hipStream_t st, int I, int n, uint64_t seed, uint32_t inst_base, double *q);
// q[s] = dense uniform instance inst_base + s (upper triangle incl. diagonal U[-1,1), mirrored)
void launch_gen_uniform(hipStream_t st, int I, int n, uint64_t seed, uint32_t inst_base,
                        double *q);
// bf16 hi+lo operand packing and fp32 diagonal sub-blocks of I instances
void launch_pack_q(hipStream_t st, int I, int n, int np, const double *q, uint16_t *qpack,
                   float *qdb);

// anneal_csr.hip
// Per-instance descriptor of a sparse batch (offsets into the batch arrays).
struct CsrInst {
  int32_t n;          // spins
  int32_t words;      // ceil((n + 1) / 32) state words (bit n: the ELL padding target)
  int32_t width;      // ELL row width (multiple of 4)
  int32_t ncol;       // colour classes (site order 1)
  int64_t ell_off;    // first ELL entry (multiple of 4)
  int64_t spin_off;   // first spin in diag / diag64
  int64_t row_off;    // first entry in row_ptr (n + 1 entries per instance)
  int64_t csr_off;    // first entry in col / val
  int64_t state_off;  // first word of the instance's [words][Rp] state planes
  int64_t rep_off;    // first replica in per-replica arrays (Rp per instance)
  int64_t class_off;  // first entry in the class starts (ncol + 1 per instance)
  int32_t sites;      // colour-class kernel: sites per wave step (4, 8 or 16)
  int32_t pad_;
};
// The device arrays of a sparse batch, for kernels outside the annealer (the
// warm-start network): instance descriptors, CSR (both directions, no
// diagonal) and the diagonal, with spin j of instance i at spin_off + j.
struct CsrView {
  int num_instances = 0;
  int max_n = 0;
  int64_t total_spins = 0;
  const CsrInst *inst = nullptr;
  const int32_t *row_ptr = nullptr;
  const int32_t *col = nullptr;
  const double *val = nullptr;
  const double *diag = nullptr;
};
// Parameters of one batched run, read by the kernel from device memory so a
// captured graph replays with new values.
struct CsrRun {
  int32_t num_items;  // work items (instance, 64-replica group)
  int32_t num_sweeps;
  int32_t Rp;         // replicas per instance, padded to a multiple of 64
  int32_t init_mode;  // 0: Philox start states, 1: caller's planes
  uint32_t k0, k1;    // Philox key
  uint32_t instance_base;
  int32_t sweep_base;  // number of the launch's first sweep (thresholds, tscale index)
  int32_t use_tsr;     // 1: replica r sweeps at tsr[rep_off + r] (replica exchange)
  int32_t pad_;
};
int csr_lds_bytes(int waves, int max_words);
// pf: the row-prefetching kernel (batches within csr_pf_eligible)
bool csr_pf_eligible(int max_width);
int csr_max_groups_per_cu(bool pf, int waves, int max_words);
void launch_csr_queue_reset(hipStream_t st, int *queue);
void launch_csr_sweeps(hipStream_t st, bool pf, int waves, int groups, int max_words,
                       const CsrInst *inst, const int2 *items, const CsrRun *run,
                       const int4 *ell_col, const float4 *ell_val, const float *diag,
                       const float *tscale, const float *tsr, const uint32_t *init_bits,
                       uint32_t *best_bits, uint32_t *final_bits, double *e_best,
                       double *e_final, int *queue);
// colour-class sweeps: instance i updates G_i = CsrInst::sites (4, 8 or 16)
// sites of 64/G_i replicas per wave step; items are (instance, group of
// 64/G_i replicas); corder lists the sites of each instance class by class
// (offset spin_off), cstart the class starts (offset class_off); pell_col /
// pell_val / pdiag hold the ELL rows and the diagonal in that visiting order
// (offsets ell_off, spin_off).  plane_words: LDS words per wave for the
// state planes, the batch's largest words_i * 64/G_i.
int csr_col_lds_bytes(int waves, int plane_words);
int csr_col_max_groups_per_cu(int waves, int plane_words);
void launch_csr_col_sweeps(hipStream_t st, int waves, int groups, int plane_words,
                           const CsrInst *inst, const int2 *items, const CsrRun *run,
                           const int4 *ell_col, const float4 *ell_val, const float *diag,
                           const int32_t *corder, const int32_t *cstart, const int4 *pell_col,
                           const float4 *pell_val, const float *pdiag, const float *tscale,
                           const float *tsr, const uint32_t *init_bits, uint32_t *best_bits,
                           uint32_t *final_bits, double *e_best, double *e_final, int *queue);
// start planes of every instance from marginals p (Σn floats, instance i at
// spin_off): the dense warm-start rule, stream instance_base + i
void launch_csr_warm_init(hipStream_t st, const CsrInst *inst, int num_instances, int R, int Rp,
                          uint64_t seed, uint32_t instance_base, const float *p,
                          uint32_t *planes);
void launch_csr_energy(hipStream_t st, const CsrInst *inst, const int2 *items, int num_items,
                       int Rp, int max_words, const int32_t *row_ptr, const int32_t *col,
                       const double *val, const double *diag64, const uint32_t *bits,
                       double *out);
// population annealing over the 64-replica items: replica r < R of every
// instance keeps the step's best planes and energy (eb, best) where they are
// below its best so far (or all of them when `first`); the gather gives
// replica j < R the final planes of replica anc[instance][j]
void launch_csr_pop_best(hipStream_t st, const CsrInst *inst, const int2 *items, int num_items,
                         int Rp, bool first, const double *eb, const uint32_t *best,
                         double *pop_eb, uint32_t *pop_best);
// replica exchange round on the GPU (exchange_round_kernel, dense and
// sparse paths): rep_at / slot_of per instance [R], ts_slots [K], e and tsr
// indexed by row i * Rp + r
void launch_exchange_round(hipStream_t st, int I, int L, int K, int R, int Rp, uint32_t round,
                           uint64_t seed, uint32_t inst_base, const float *ts_slots,
                           const double *e, uint32_t *rep_at, uint32_t *slot_of, float *tsr);
void launch_csr_pop_gather(hipStream_t st, const CsrInst *inst, const int2 *items, int num_items,
                           int Rp, int R, const uint32_t *anc, const uint32_t *fin,
                           uint32_t *init);

}  // namespace os
