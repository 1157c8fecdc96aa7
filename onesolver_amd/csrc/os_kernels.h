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
void launch_energy_fp64(hipStream_t st, int n, int W, const double *q, int count,
                        const uint32_t *bits, double *out, double *scratch, float *fields, int ldf,
                        unsigned int *drift_max);

// anneal_ref.hip
int ref_lds_limit();
size_t ref_lds_bytes(int n);
void launch_ref_anneal(hipStream_t st, int n, const double *q, const double *beta, int num_iter,
                       int spb, uint32_t num_tries, uint64_t seed, double *gfields,
                       uint32_t *best_bits, double *best_run);

// anneal_dense.hip
int dense_max_n();
int dense_padded_n(int n);
void launch_dense_sweeps(hipStream_t st, int n, int np, const uint4 *qpack, const float *qdb,
                         const float *tscale, int sweep0, int nsweeps, uint64_t seed,
                         uint32_t instance, int R_pad, float *fields, uint32_t *xbits,
                         uint32_t *bestbits, double *energy, double *ebest, int *status);
void launch_sweep_init(hipStream_t st, int n, int npw, int R, uint64_t seed, uint32_t instance,
                       uint32_t *xbits);
void launch_warm_init(hipStream_t st, int n, int npw, int R, uint64_t seed, uint32_t instance,
                      const float *p, uint32_t *xbits);

}  // namespace os
