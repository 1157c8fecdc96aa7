// anneal_dense.hip -- sweep-mode simulated annealing of dense QUBOs with
// delayed field updates on MFMA (gfx950).
//
// Replicas run Metropolis sweeps in a fixed site order j = 0..N-1.  Local
// fields h_rj = Q_jj + sum_{l!=j} Q_jl x_rl live in MFMA accumulators (fp32).
// Spins are decided in blocks of k = 16; after block b the flip deltas D_b
// (replicas x 16, values -1/0/+1, exact in bf16) update every field:
// F += [D D] * [Qhi; Qlo] on v_mfma_f32_16x16x32_bf16 (Q = Qhi + Qlo; the
// K=32 operand pairs D with the hi and with the lo half of Q's rows).
//
// Work-group: 64 replicas, 4 waves (one per SIMD, so each wave has the full
// register file and the fields sit in accumulation registers); wave w owns
// the TPW column tiles (16 spins each) [w*TPW, (w+1)*TPW).  Block b belongs
// to the wave owning its columns.  For block g of the global block sequence
// (g = sweep * NB + b) the owner, lane = replica:
//   (a) applies D_{g-1} to block b's tile (4 MFMAs), transposes it through
//       LDS and decides its 16 spins in order; flips inside the block correct
//       the later spins with the fp32 16x16 diagonal sub-block of Q;
//   (b) applies D_{g-1} to its other tiles, one tile per decision step, so the
//       matrix cores run under the decisions' VALU work;
//   (c) publishes D_g as +/- bit masks in an LDS ring and bumps a counter.
// The owner's Q operands come from LDS only: the three other waves load them
// (a third each) early in the previous step and store them in a double-
// buffered LDS area at its end.  Those waves also apply each published D
// block to all their own tiles (their operands are prefetched a full block
// ahead in registers) and compute the acceptance thresholds two blocks ahead.  Waves synchronise only through LDS counters
// with bounded spin waits (an expired wait aborts the launch and reports it
// in `status`).
//
// Acceptance: flip iff dE < neg_ln_u(w) * tscale[t] with w the Philox word of
// counter (j>>2, replica, sweep, tag), i.e. with probability
// min(1, exp(-dE / tscale^-1)).
#include "os_kernels.h"

namespace os {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

namespace {

constexpr int kWaves = 4;
constexpr int kRT = 64;    // replicas per work-group
constexpr int kRing = 32;  // D blocks kept in LDS
constexpr int kTRing = 4;  // threshold blocks kept in LDS
constexpr int kLA = 2;     // thresholds are computed this many blocks ahead
constexpr int kHelpers = kWaves - 1;
constexpr int kSpinLimit = 1 << 24;

__device__ __forceinline__ bf16x8 as_bf16x8(uint4 v) { return __builtin_bit_cast(bf16x8, v); }

// Synchronisation between the waves only concerns LDS data written with
// ordinary ds_write instructions.  LDS operations of one wave complete in
// order, so publishing needs only "all my LDS writes done" (lgkmcnt(0)) and a
// compiler barrier -- not the vmcnt(0) wait that a workgroup release fence
// implies, which would stall on this wave's outstanding operand loads.
__device__ __forceinline__ void lds_writes_done() {
  __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0), vmcnt/expcnt unconstrained
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

__device__ __forceinline__ void wave_sync() {
  lds_writes_done();
  __builtin_amdgcn_wave_barrier();
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

__device__ __forceinline__ int lds_load(int *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Publishes v in *p after all of this wave's earlier LDS writes.
__device__ __forceinline__ void lds_store(int *p, int v) {
  lds_writes_done();
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Waits until *p >= target, then orders later LDS reads after the wait.
// Returns false (and raises *abort) if the wait exceeds kSpinLimit sleeps or
// another wave has aborted.
__device__ __forceinline__ bool wait_ge(int *p, int target, int *abort) {
  int spins = 0;
  while (lds_load(p) < target) {
    if (lds_load(abort)) return false;
    __builtin_amdgcn_s_sleep(1);
    if (++spins > kSpinLimit) {
      lds_store(abort, 1);
      return false;
    }
  }
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
  return true;
}

// thresholds of spins [16*bn + 4q, +4) of replica r for sweep (sweep0+sw)
__device__ __forceinline__ void block_thresholds(int item, int bn, int sw, int n, uint32_t rbase,
                                                 int sweep0, uint32_t k0, uint32_t k1,
                                                 uint32_t tag, float ts, float (*thr)[17]) {
  const int r = item >> 2, q = item & 3;
  const int j0 = bn * 16 + 4 * q;
  const uint4 o = philox4x32_10(make_uint4((uint32_t)(j0 >> 2), rbase + (uint32_t)r,
                                           (uint32_t)(sweep0 + sw), tag),
                                k0, k1);
#pragma unroll
  for (int jj = 0; jj < 4; ++jj) {
    const int j = j0 + jj;
    thr[r][4 * q + jj] = j < n ? neg_ln_u(pick(o, jj)) * ts : -INFINITY;
  }
}

// Q operand (16 bytes per lane) of block `blk`, this wave's column tile tt.
__device__ __forceinline__ uint4 q_operand(const uint4 *__restrict__ qpack, int np, int blk,
                                           int col0, int tt, int lrow, int kg) {
  return qpack[((size_t)blk * np + col0 + 16 * tt + lrow) * 4 + kg];
}

// A operands of a D block: replica 16m + lrow, spins 8*(kg&1) .. +7, from the
// ring's +/- masks (low 16 bits: flips 0->1, high 16 bits: flips 1->0)
// through the 8-bit -> 8 x bf16 tables (1.0 resp. -1.0 where the bit is set).
__device__ __forceinline__ void load_a(const uint32_t *slot, const uint4 *lutp, const uint4 *lutn,
                                       int lrow, int kg, bf16x8 (&A)[4]) {
  const int sh = 8 * (kg & 1);
#pragma unroll
  for (int m = 0; m < 4; ++m) {
    const uint32_t wd = slot[16 * m + lrow];
    const uint4 p = lutp[(wd >> sh) & 0xFFu];
    const uint4 q = lutn[(wd >> (16 + sh)) & 0xFFu];
    A[m] = as_bf16x8(make_uint4(p.x | q.x, p.y | q.y, p.z | q.z, p.w | q.w));
  }
}

// Asynchronous copy of block `blk`'s diagonal 16x16 sub-block (1 KB) into LDS.
__device__ __forceinline__ void stage_qd(const float *__restrict__ qdb, int blk, int lane,
                                         uint4 *dst) {
  __builtin_amdgcn_global_load_lds(
      (__attribute__((address_space(1))) void *)(reinterpret_cast<const uint4 *>(qdb) + (size_t)blk * 64 + lane),
      (__attribute__((address_space(3))) void *)dst, 16, 0, 0);
}

// Asynchronous copy (global_load_lds_dwordx4) of the Q operands of block
// `blk` for the column tiles [t0, t1) of the wave whose columns start at col0:
// lane l's 16 bytes land at dst[tt][l], so a lane later reads back exactly the
// operand it would have loaded.
__device__ __forceinline__ void stage_tiles(const uint4 *__restrict__ qpack, int np, int blk,
                                            int col0, int t0, int t1, int lrow, int kg,
                                            uint4 (*dst)[64]) {
  for (int tt = t0; tt < t1; ++tt)
    __builtin_amdgcn_global_load_lds(
        (__attribute__((address_space(1))) void *)(qpack + ((size_t)blk * np + col0 + 16 * tt + lrow) * 4 + kg),
        (__attribute__((address_space(3))) void *)&dst[tt][0], 16, 0, 0);
}

// Non-owner step: wait for D_next and apply it to every tile with the
// operands staged in this wave's LDS buffer (block sb_blk, staged now if
// stale), then stage the following block's operands, which land while the
// wave waits for the next D block.  Returns false if the wait expired.
template <int TPW>
__device__ __forceinline__ bool follow_one(const uint4 *__restrict__ qpack, int np, int col0,
                                           int lrow, int kg, int lane, int w, int G,
                                           const uint32_t (*dring)[kRT], const uint4 *lutp,
                                           const uint4 *lutn, int *d_count, int *wnext,
                                           int *abort_flag, int &next, int &sb_blk,
                                           uint4 (*sb)[64], f32x4 (&C)[4][TPW]) {
  const int nb = np / 16;
  if (sb_blk != next % nb) {
    stage_tiles(qpack, np, next % nb, col0, 0, TPW, lrow, kg, sb);
    sb_blk = next % nb;
  }
  if (!wait_ge(d_count, next + 1, abort_flag)) return false;
  bf16x8 A[4];
  load_a(dring[next % kRing], lutp, lutn, lrow, kg, A);
  __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0): the staged operands have landed
#pragma unroll
  for (int tt = 0; tt < TPW; ++tt) {
    const bf16x8 Bt = as_bf16x8(sb[tt][lane]);
#pragma unroll
    for (int m = 0; m < 4; ++m)
      C[m][tt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[m], Bt, C[m][tt], 0, 0, 0);
  }
  ++next;
  if (lane == 0) lds_store(&wnext[w], next);  // also waits for the reads of sb above
  if (next < G) {
    stage_tiles(qpack, np, next % nb, col0, 0, TPW, lrow, kg, sb);
    sb_blk = next % nb;
  }
  return true;
}

}  // namespace

template <int TPW>
__global__ __launch_bounds__(256, 1) void dense_sweep_kernel(
    int n, const uint4 *__restrict__ qpack, const float *__restrict__ qdb,
    const float *__restrict__ tscale, int sweep0, int nsweeps, uint32_t k0, uint32_t k1,
    uint32_t tag, float *__restrict__ fields, uint32_t *__restrict__ xbits,
    uint32_t *__restrict__ bestbits, double *__restrict__ energy, double *__restrict__ ebest,
    int *__restrict__ status) {
  static_assert(TPW <= 16, "one interleaved tile per decision step");
  constexpr int NP = kWaves * 16 * TPW;
  constexpr int NB = NP / 16;
  constexpr int NPW = NP / 32;
  __shared__ uint32_t dring[kRing][kRT];
  __shared__ uint4 lutp[256];
  __shared__ uint4 lutn[256];
  __shared__ float fblk[kRT][17];
  __shared__ float thr[kTRing][kRT][17];
  __shared__ uint32_t xs[kRT][NPW + 1];
  __shared__ double es[kRT];
  __shared__ double ebs[kRT];
  __shared__ __align__(16) uint4 qds[2][64];  // staged diagonal 16x16 sub-block of Q (fp32)
  __shared__ __align__(16) uint4 ob[2][TPW][64];   // owner's Q operands, staged by the helpers
  __shared__ __align__(16) uint4 sb[kWaves][TPW][64];  // each wave's own staged operands
  __shared__ int d_count;
  __shared__ int ob_cnt[2];
  __shared__ int thr_cnt[kTRing];
  __shared__ int wnext[kWaves];
  __shared__ int abort_flag;

  const int tid = threadIdx.x;
  const int w = tid >> 6;
  const int lane = tid & 63;
  const int lrow = lane & 15;
  const int kg = lane >> 4;
  const uint32_t rbase = blockIdx.x * kRT;
  const int col0 = w * 16 * TPW;
  const int G = nsweeps * NB;

  for (int i = tid; i < kRT * NPW; i += 256)
    xs[i / NPW][i % NPW] = xbits[(size_t)(rbase + i / NPW) * NPW + i % NPW];
  {
    uint32_t v[4], u[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const uint32_t lo = (tid >> (2 * q)) & 1u, hi = (tid >> (2 * q + 1)) & 1u;
      v[q] = (lo ? 0x3F80u : 0u) | (hi ? 0x3F800000u : 0u);
      u[q] = (lo ? 0xBF80u : 0u) | (hi ? 0xBF800000u : 0u);
    }
    lutp[tid] = make_uint4(v[0], v[1], v[2], v[3]);
    lutn[tid] = make_uint4(u[0], u[1], u[2], u[3]);
  }
  if (tid < kRT) {
    es[tid] = energy[rbase + tid];
    ebs[tid] = ebest[rbase + tid];
  }
  if (tid < kTRing) thr_cnt[tid] = tid < kLA ? kHelpers : 0;
  if (tid < kWaves) wnext[tid] = 0;
  if (tid < 2) ob_cnt[tid] = 0;
  if (tid == 0) {
    d_count = 0;
    abort_flag = 0;
  }
  f32x4 C[4][TPW];
#pragma unroll
  for (int m = 0; m < 4; ++m)
#pragma unroll
    for (int tt = 0; tt < TPW; ++tt)
#pragma unroll
      for (int i = 0; i < 4; ++i)
        C[m][tt][i] =
            fields[(size_t)(rbase + 16 * m + 4 * kg + i) * NP + col0 + 16 * tt + lrow];
  // thresholds of the first kLA blocks (256 items each)
  for (int it = tid; it < kLA * 256; it += 256) {
    const int gblk = it >> 8;
    if (gblk < G)
      block_thresholds(it & 255, gblk % NB, gblk / NB, n, rbase, sweep0, k0, k1, tag,
                       tscale[gblk / NB], thr[gblk]);
  }
  __syncthreads();

  int next = 0;  // D blocks applied to all of this wave's tiles
  bool ok = true;
  int sb_blk = 0;  // block whose operands this wave's sb buffer holds
  stage_tiles(qpack, NP, 0, col0, 0, TPW, lrow, kg, sb[w]);
  if (w == 0) stage_qd(qdb, 0, lane, qds[0]);
  for (int g = 0; g < G && ok; ++g) {
    const int b = g % NB;
    const int owner = b / TPW;
    const int owner1 = ((g + 1) % NB) / TPW;  // owner of step g+1
    if (owner == w) {
      const int i = b % TPW;  // this wave's tile holding block b
      bf16x8 A[4];
      const uint4(*bt)[64] = ob[g & 1];  // D_{g-1}'s operands for this wave's tiles
      if (g > 0) {
        ok = wait_ge(&d_count, g, &abort_flag) &&
             wait_ge(&ob_cnt[g & 1], kHelpers * ((g + 1) >> 1), &abort_flag);
        if (!ok) break;
        load_a(dring[(g - 1) % kRing], lutp, lutn, lrow, kg, A);
      }
      if (g == 0) __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0): qds[0] staged above
      // (a) bring tile i up to date and transpose it (static indices, uniform branch)
#pragma unroll
      for (int tt = 0; tt < TPW; ++tt) {
        if (tt != i) continue;
        if (g > 0) {
          const bf16x8 Bt = as_bf16x8(bt[tt][lane]);
#pragma unroll
          for (int m = 0; m < 4; ++m)
            C[m][tt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[m], Bt, C[m][tt], 0, 0, 0);
        }
#pragma unroll
        for (int m = 0; m < 4; ++m)
#pragma unroll
          for (int q = 0; q < 4; ++q) fblk[16 * m + 4 * kg + q][lrow] = C[m][tt][q];
      }
      wave_sync();
      ok = wait_ge(&thr_cnt[g % kTRing], kHelpers * (g / kTRing + 1), &abort_flag);
      if (!ok) break;
      // ---- sequential decisions of block b, lane = replica ----
      const float(*tb_thr)[17] = thr[g % kTRing];
      float c[16];
#pragma unroll
      for (int jj = 0; jj < 16; ++jj) c[jj] = 0.0f;
      const uint32_t word = xs[lane][b >> 1];
      const int sh = (b & 1) * 16;
      uint32_t xb = (word >> sh) & 0xFFFFu;
      const uint32_t x0 = xb;
      float desum = 0.0f;
      // Software-pipelined: step jj+1's field, threshold, sub-block row (row ==
      // column: the diagonal block of a symmetric Q is symmetric) and the
      // operand of the tile updated in step jj+1 are read from LDS during step
      // jj; sched_barrier keeps the compiler from hoisting all 16 steps' reads.
      const float4 *qdv = reinterpret_cast<const float4 *>(qds[g & 1]);
      float fn = fblk[lane][0], tn = tb_thr[lane][0];
      float4 qn[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) qn[k] = qdv[k];
      uint4 bq = bt[0][lane];
#pragma unroll
      for (int jj = 0; jj < 16; ++jj) {
        const float fc = fn, tc = tn;
        const uint4 bcur = bq;
        if (jj + 1 < TPW) bq = bt[jj + 1][lane];
        float qc[16];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          qc[4 * k] = qn[k].x;
          qc[4 * k + 1] = qn[k].y;
          qc[4 * k + 2] = qn[k].z;
          qc[4 * k + 3] = qn[k].w;
        }
        if (jj < 15) {
          fn = fblk[lane][jj + 1];
          tn = tb_thr[lane][jj + 1];
#pragma unroll
          for (int k = 0; k < 4; ++k)
            if (4 * k + 3 > jj + 1) qn[k] = qdv[(jj + 1) * 4 + k];
        }
        const float h = fc + c[jj];
        const bool xj = (xb >> jj) & 1u;
        const float de = xj ? -h : h;
        const bool acc = de < tc;
        const float d = acc ? (xj ? -1.0f : 1.0f) : 0.0f;
        xb ^= acc ? (1u << jj) : 0u;
        desum += acc ? de : 0.0f;
#pragma unroll
        for (int mm = jj + 1; mm < 16; ++mm) c[mm] = fmaf(qc[mm], d, c[mm]);
        // (b) D_{g-1} into tile jj of this wave (tile i already has it)
        if (jj < TPW && g > 0 && jj != i) {
          const bf16x8 Bt = as_bf16x8(bcur);
#pragma unroll
          for (int tt = 0; tt < TPW; ++tt) {
            if (tt != jj) continue;
#pragma unroll
            for (int m = 0; m < 4; ++m)
              C[m][tt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[m], Bt, C[m][tt], 0, 0, 0);
          }
        }
        __builtin_amdgcn_sched_barrier(0);
      }
      xs[lane][b >> 1] = (word & ~(0xFFFFu << sh)) | (xb << sh);
      const uint32_t flips = xb ^ x0;
      const uint32_t dmask = (flips & ~x0) | ((flips & x0) << 16);
      // ring slot g % kRing: every other wave must have consumed D_{g - kRing}
      if (g >= kRing) {
        for (int v = 0; v < kWaves && ok; ++v)
          if (v != w) ok = wait_ge(&wnext[v], g - kRing + 1, &abort_flag);
        if (!ok) break;
      }
      dring[g % kRing][lane] = dmask;
      const double e_now = es[lane] + (double)desum;
      es[lane] = e_now;
      // end of a sweep: best-state tracking (before the next owner may change xs)
      if (b == NB - 1 && e_now < ebs[lane]) {
        ebs[lane] = e_now;
        uint32_t *dst = bestbits + (size_t)(rbase + lane) * NPW;
        for (int q = 0; q < NPW; ++q) dst[q] = xs[lane][q];
      }
      next = g;  // every tile now holds D_0 .. D_{g-1}
      if (lane == 0) {
        lds_store(&d_count, g + 1);
        lds_store(&wnext[w], next);
      }
      if (owner1 != w && g + 1 < G) {
        // end of the turn: this wave applies D_g next, as a non-owner
        stage_tiles(qpack, NP, b, col0, 0, TPW, lrow, kg, sb[w]);
        sb_blk = b;
      }
    } else {
      const int hw = (w - owner - 1 + kWaves) % kWaves;  // 0 .. kHelpers-1
      const bool staging = g + 1 < G;
      const int lo = hw * TPW / kHelpers, hi = (hw + 1) * TPW / kHelpers;
      // (2) thresholds kLA blocks ahead: 256 items, the extra 64 rotate
      if (g + kLA < G) {
        const int g2 = g + kLA;
        block_thresholds(hw * 64 + lane, g2 % NB, g2 / NB, n, rbase, sweep0, k0, k1, tag,
                         tscale[g2 / NB], thr[g2 % kTRing]);
        if (hw == g % kHelpers)
          block_thresholds(192 + lane, g2 % NB, g2 / NB, n, rbase, sweep0, k0, k1, tag,
                           tscale[g2 / NB], thr[g2 % kTRing]);
        lds_writes_done();
        if (lane == 0)
          __hip_atomic_fetch_add(&thr_cnt[g2 % kTRing], 1, __ATOMIC_RELAXED,
                                 __HIP_MEMORY_SCOPE_WORKGROUP);
      }
      // (3) follow the D stream up to D_{g-1}
      while (next < g) {
        ok = follow_one<TPW>(qpack, NP, col0, lrow, kg, lane, w, G, dring, lutp, lutn, &d_count,
                             wnext, &abort_flag, next, sb_blk, sb[w], C);
        if (!ok) break;
      }
      // (4) stage this wave's share of the next owner's operands (block b) in
      //     ob[(g+1)&1], last read in step g-1, which is complete (this wave
      //     has seen D_{g-1}); the count is raised once the copies have landed
      if (staging && ok) {
        stage_tiles(qpack, NP, b, owner1 * 16 * TPW, lo, hi, lrow, kg, ob[(g + 1) & 1]);
        if (hw == 0) stage_qd(qdb, (g + 1) % NB, lane, qds[(g + 1) & 1]);
        __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0)
        if (lane == 0)
          __hip_atomic_fetch_add(&ob_cnt[(g + 1) & 1], 1, __ATOMIC_RELAXED,
                                 __HIP_MEMORY_SCOPE_WORKGROUP);
      }
    }
  }
  // drain: apply the remaining D blocks
  while (ok && next < G)
    ok = follow_one<TPW>(qpack, NP, col0, lrow, kg, lane, w, G, dring, lutp, lutn, &d_count, wnext,
                         &abort_flag, next, sb_blk, sb[w], C);
  if (!ok && lane == 0) lds_store(&abort_flag, 1);
#pragma unroll
  for (int m = 0; m < 4; ++m)
#pragma unroll
    for (int tt = 0; tt < TPW; ++tt)
#pragma unroll
      for (int i = 0; i < 4; ++i)
        fields[(size_t)(rbase + 16 * m + 4 * kg + i) * NP + col0 + 16 * tt + lrow] = C[m][tt][i];
  __syncthreads();
  for (int i = tid; i < kRT * NPW; i += 256)
    xbits[(size_t)(rbase + i / NPW) * NPW + i % NPW] = xs[i / NPW][i % NPW];
  if (tid < kRT) {
    energy[rbase + tid] = es[tid];
    ebest[rbase + tid] = ebs[tid];
  }
  if (tid == 0 && abort_flag) atomicOr(status, 1);
}

// Random initial bits: spin j of replica r from counter (j>>7, r, 0, tag_init).
__global__ __launch_bounds__(256) void sweep_init_kernel(int n, int npw, int R, uint32_t k0,
                                                         uint32_t k1, uint32_t tag,
                                                         uint32_t *__restrict__ xbits) {
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (size_t)R * npw) return;
  const uint32_t r = (uint32_t)(idx / npw);
  const int wd = (int)(idx % npw);
  const int j0 = wd * 32;
  const uint4 o = philox4x32_10(make_uint4((uint32_t)(j0 >> 7), r, 0u, tag), k0, k1);
  uint32_t v = pick(o, (j0 >> 5) & 3);
  const int valid = n - j0;
  if (valid < 32) v = valid <= 0 ? 0u : (v & ((1u << valid) - 1u));
  xbits[idx] = v;
}

// Warm start: x_j = (u < p_j), u = (w >> 8) * 2^-24 from counter (j>>2, r, 0, tag_warm).
__global__ __launch_bounds__(256) void warm_init_kernel(int n, int npw, int R, uint32_t k0,
                                                        uint32_t k1, uint32_t tag,
                                                        const float *__restrict__ p,
                                                        uint32_t *__restrict__ xbits) {
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (size_t)R * npw) return;
  const uint32_t r = (uint32_t)(idx / npw);
  const int wd = (int)(idx % npw);
  uint32_t v = 0;
  for (int q = 0; q < 8; ++q) {
    const int j0 = wd * 32 + 4 * q;
    if (j0 >= n) break;
    const uint4 o = philox4x32_10(make_uint4((uint32_t)(j0 >> 2), r, 0u, tag), k0, k1);
    for (int jj = 0; jj < 4; ++jj) {
      const int j = j0 + jj;
      if (j < n && (float)(pick(o, jj) >> 8) * 0x1p-24f < p[j]) v |= 1u << (4 * q + jj);
    }
  }
  xbits[idx] = v;
}

int dense_max_n() { return kWaves * 16 * 16; }

// Tiles per wave are instantiated for these counts; n is padded up to the
// next one (64 columns per tile count step).
static const int kTpwList[] = {1, 2, 4, 6, 8, 12, 16};

int dense_padded_n(int n) {
  const int need = (n + kWaves * 16 - 1) / (kWaves * 16);
  for (int t : kTpwList)
    if (t >= need) return t * kWaves * 16;
  return -1;
}

void launch_dense_sweeps(hipStream_t st, int n, int np, const uint4 *qpack, const float *qdb,
                         const float *tscale, int sweep0, int nsweeps, uint64_t seed,
                         uint32_t instance, int R_pad, float *fields, uint32_t *xbits,
                         uint32_t *bestbits, double *energy, double *ebest, int *status) {
  const uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
  const uint32_t tag = (instance << 8) | TAG_SWEEP_THR;
  const dim3 grid(R_pad / kRT), block(kWaves * 64);
  switch (np / (kWaves * 16)) {
#define OS_DENSE_CASE(T)                                                                    \
  case T:                                                                                   \
    dense_sweep_kernel<T><<<grid, block, 0, st>>>(n, qpack, qdb, tscale, sweep0, nsweeps, k0, \
                                                  k1, tag, fields, xbits, bestbits, energy, \
                                                  ebest, status);                           \
    break;
    OS_DENSE_CASE(1)
    OS_DENSE_CASE(2)
    OS_DENSE_CASE(4)
    OS_DENSE_CASE(6)
    OS_DENSE_CASE(8)
    OS_DENSE_CASE(12)
    OS_DENSE_CASE(16)
#undef OS_DENSE_CASE
    default:
      throw Error(-3, "dense sweep kernel: unsupported padded size");
  }
  OS_HIP_CHECK(hipGetLastError());
}

void launch_sweep_init(hipStream_t st, int n, int npw, int R, uint64_t seed, uint32_t instance,
                       uint32_t *xbits) {
  const size_t total = (size_t)R * npw;
  sweep_init_kernel<<<(unsigned)((total + 255) / 256), 256, 0, st>>>(
      n, npw, R, (uint32_t)seed, (uint32_t)(seed >> 32), (instance << 8) | TAG_SWEEP_INIT, xbits);
  OS_HIP_CHECK(hipGetLastError());
}

void launch_warm_init(hipStream_t st, int n, int npw, int R, uint64_t seed, uint32_t instance,
                      const float *p, uint32_t *xbits) {
  const size_t total = (size_t)R * npw;
  warm_init_kernel<<<(unsigned)((total + 255) / 256), 256, 0, st>>>(
      n, npw, R, (uint32_t)seed, (uint32_t)(seed >> 32), (instance << 8) | TAG_WARM, p, xbits);
  OS_HIP_CHECK(hipGetLastError());
}

}  // namespace os
