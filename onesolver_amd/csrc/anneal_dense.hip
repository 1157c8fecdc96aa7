// anneal_dense.hip -- sweep-mode simulated annealing of dense QUBOs with
// delayed field updates on MFMA (gfx950).
//
// Replicas run Metropolis sweeps in a fixed site order j = 0..N-1.  Local
// fields h_rj = Q_jj + sum_{l!=j} Q_jl x_rl live in MFMA accumulators (fp32).
// Spins are decided in blocks of k = 16; after block b the flip deltas D_b
// (replicas x 16, values -1/0/+1, exact in bf16) update every field:
// F += D * Qhi + D * Qlo on v_mfma_f32_32x32x16_bf16 (Q = Qhi + Qlo, both
// bf16, rows 16b .. 16b+15 of Q).
//
// Work-group: 64 replicas (two 32-replica tiles), 4 waves (one per SIMD, so
// each wave has the full register file and the fields sit in accumulation
// registers).  Blocks are dealt to the waves round-robin: block b belongs to
// wave b % 4, and wave w holds the columns of its blocks as TPW tiles of 32
// columns (tile t = blocks w + 8t and w + 8t + 4).  Step g of the global
// block sequence (g = sweep * NB + b):
//   * the owner of block b applies D_{g-1} to the one tile holding b (4
//     MFMAs), transposes it through LDS, decides the 16 spins in order (lane
//     = replica; flips inside the block correct the later spins with the fp32
//     16x16 diagonal sub-block of Q) and publishes D_g as +/- bit masks in an
//     LDS ring; only then does it apply D_{g-1} to its other tiles;
//   * the three other waves apply D_{g-1} to all their tiles, with the
//     Philox rounds of the acceptance thresholds kLA blocks ahead spread
//     between the MFMAs.
// So the chain of decisions passes from SIMD to SIMD, and each SIMD's MFMA
// work runs while the next owner decides.  Q operands stream from L2 into
// registers one step ahead (the owner's tile and sub-block two steps ahead);
// waves synchronise only through LDS counters with bounded spin waits (an
// expired wait aborts the launch and reports it in `status`).
//
// Acceptance: flip iff dE < neg_ln_u(w) * tscale[t] with w the Philox word of
// counter (j>>2, replica, sweep, tag), i.e. with probability
// min(1, exp(-dE / tscale^-1)).  With `tsr` (replica exchange) the scale is
// tsr[row] of the replica instead, for the whole launch.
#include "os_kernels.h"

namespace os {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

namespace {

constexpr int kWaves = 4;
constexpr int kRT = 64;    // replicas per work-group
constexpr int kRing = 8;   // D blocks kept in LDS
constexpr int kTRing = 4;  // threshold blocks kept in LDS
constexpr int kLA = 3;     // thresholds are computed this many blocks ahead
constexpr int kHelpers = kWaves - 1;
// The last kThrWriters helpers of a step compute the thresholds of a later
// block.  The first helpers own the next steps: their work in this step delays
// the decision chain.
constexpr int kThrWriters = 1;
constexpr int kThrItems = 4 / kThrWriters;  // threshold items per lane and writer
static_assert(kThrWriters >= 1 && kThrWriters <= kHelpers && 4 % kThrWriters == 0, "writers");
constexpr int kFT = kRT + 4;  // row stride (floats) of the transposed field tile
constexpr int kSpinLimit = 1 << 24;

static_assert(kLA < kTRing, "a threshold slot is rewritten only after its block was decided");
// The owner of step g writes ring slot g % kRing after it has seen D_{g-1}
// published.  The owners of steps g-4 .. g-1 -- all four waves -- published
// only after their own load_a of those steps, so every wave has read
// D_{g-5} and older; slot g % kRing last held D_{g-kRing}.
static_assert(kRing >= kWaves + 1, "a ring slot is rewritten only after every wave read it");

__device__ __forceinline__ bf16x8 as_bf16x8(uint4 v) { return __builtin_bit_cast(bf16x8, v); }

__device__ __forceinline__ f32x16 mfma32(bf16x8 a, bf16x8 b, f32x16 c) {
  return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
}

// Synchronisation between the waves only concerns LDS data written with
// ordinary ds_write instructions.  LDS operations of one wave complete in
// order, so publishing needs only "all my LDS writes done" (lgkmcnt(0)) and a
// compiler barrier -- not the vmcnt(0) wait that a workgroup release fence
// implies, which would stall on this wave's outstanding operand loads.  The
// counters are accessed by ds instructions in asm, so no memory-model fence
// or atomic lowering adds such a wait either; each asm load waits for its own
// result (lgkmcnt(0)) inside the statement.
__device__ __forceinline__ void lds_writes_done() {
  __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0), vmcnt/expcnt unconstrained
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

__device__ __forceinline__ void wave_sync() {
  lds_writes_done();
  __builtin_amdgcn_wave_barrier();
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

__device__ __forceinline__ uint32_t lds_addr(const void *p) {
  return (uint32_t)(uintptr_t)(const __attribute__((address_space(3))) void *)p;
}

__device__ __forceinline__ int lds_load(const int *p) {
  int v;
  asm volatile("ds_read_b32 %0, %1\n\ts_waitcnt lgkmcnt(0)" : "=v"(v) : "v"(lds_addr(p)) : "memory");
  return v;
}

// Publishes v in *p after all of this wave's earlier LDS writes.
__device__ __forceinline__ void lds_store(int *p, int v) {
  asm volatile("s_waitcnt lgkmcnt(0)\n\tds_write_b32 %0, %1" : : "v"(lds_addr(p)), "v"(v) : "memory");
}

__device__ __forceinline__ void lds_add(int *p) {
  asm volatile("s_waitcnt lgkmcnt(0)\n\tds_add_u32 %0, %1" : : "v"(lds_addr(p)), "v"(1) : "memory");
}

// Waits until *p >= target, then orders later LDS reads after the wait.
// Returns false (and raises *abort) if the wait exceeds kSpinLimit sleeps or
// another wave has aborted (checked every 64 sleeps).
__device__ __forceinline__ bool wait_ge(int *p, int target, int *abort) {
  int spins = 0;
  while (lds_load(p) < target) {
    __builtin_amdgcn_s_sleep(1);
    if ((++spins & 63) == 0) {
      if (lds_load(abort)) return false;
      if (spins > kSpinLimit) {
        lds_store(abort, 1);
        return false;
      }
    }
  }
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
  return true;
}

// thresholds of spins [16*bn + 4q, +4) of replica r = item/4 (q = item%4) for
// sweep (sweep0+sw); spins beyond n never flip
__device__ __forceinline__ void item_thresholds(int item, int bn, int sw, int n, uint32_t rbase,
                                                int sweep0, uint32_t k0, uint32_t k1,
                                                uint32_t tag, float ts, float (&v)[4]) {
  const int r = item >> 2, q = item & 3;
  const int j0 = bn * 16 + 4 * q;
  const uint4 o = philox4x32_10(make_uint4((uint32_t)(j0 >> 2), rbase + (uint32_t)r,
                                           (uint32_t)(sweep0 + sw), tag),
                                k0, k1);
#pragma unroll
  for (int jj = 0; jj < 4; ++jj) {
    const float x = neg_ln_u(pick(o, jj)) * ts;  // computed unconditionally: no branch
    v[jj] = j0 + jj < n ? x : -INFINITY;
  }
}

__device__ __forceinline__ void store_thresholds(int item, const float (&v)[4],
                                                 float (*thr)[17]) {
  const int r = item >> 2, q = item & 3;
#pragma unroll
  for (int jj = 0; jj < 4; ++jj) thr[r][4 * q + jj] = v[jj];
}

__device__ __forceinline__ void block_thresholds(int item, int bn, int sw, int n, uint32_t rbase,
                                                 int sweep0, uint32_t k0, uint32_t k1,
                                                 uint32_t tag, float ts, float (*thr)[17]) {
  float v[4];
  item_thresholds(item, bn, sw, n, rbase, sweep0, k0, k1, tag, ts, v);
  store_thresholds(item, v, thr);
}

// Round r (0-based) of Philox4x32-10; ten rounds from (ctr, k0, k1) give
// philox4x32_10(ctr, k0, k1).
__device__ __forceinline__ void philox_round(uint4 &c, uint32_t &k0, uint32_t &k1, int r) {
  if (r) {
    k0 += kPhiloxW0;
    k1 += kPhiloxW1;
  }
  const uint64_t p0 = (uint64_t)kPhiloxM0 * c.x;
  const uint64_t p1 = (uint64_t)kPhiloxM1 * c.z;
  c = make_uint4((uint32_t)(p1 >> 32) ^ c.y ^ k0, (uint32_t)p1, (uint32_t)(p0 >> 32) ^ c.w ^ k1,
                 (uint32_t)p0);
}

// A operands of a D block: replica 32m + lc, spins 8*kh .. +7, from the ring's
// +/- masks (low 16 bits: flips 0->1, high 16 bits: flips 1->0) through the
// 8-bit -> 8 x bf16 tables (1.0 resp. -1.0 where the bit is set).
__device__ __forceinline__ void load_a(const uint32_t *slot, const uint4 *lutp, const uint4 *lutn,
                                       int lc, int kh, bf16x8 (&A)[2]) {
  const int sh = 8 * kh;
#pragma unroll
  for (int m = 0; m < 2; ++m) {
    const uint32_t wd = slot[32 * m + lc];
    const uint4 p = lutp[(wd >> sh) & 0xFFu];
    const uint4 q = lutn[(wd >> (16 + sh)) & 0xFFu];
    A[m] = as_bf16x8(make_uint4(p.x | q.x, p.y | q.y, p.z | q.z, p.w | q.w));
  }
}


// Column of lane lc in tile t of wave w: the tile holds blocks w + 8t (lanes
// 0..15) and w + 8t + 4 (lanes 16..31).
__device__ __forceinline__ int tile_col(int w, int t, int lc) {
  return 16 * (w + 8 * t + 4 * (lc >> 4)) + (lc & 15);
}

}  // namespace

template <int TPW>
__global__ __launch_bounds__(256, 1) void dense_sweep_kernel(
    int n, const uint4 *__restrict__ qpack, const float *__restrict__ qdb,
    const float *__restrict__ tscale, const float *__restrict__ tsr, int sweep0, int nsweeps,
    uint32_t k0, uint32_t k1,
    uint32_t inst_base, int gpi, float *__restrict__ fields, uint32_t *__restrict__ xbits,
    uint32_t *__restrict__ bestbits, double *__restrict__ energy, double *__restrict__ ebest,
    int *__restrict__ status) {
  static_assert(TPW >= 1 && TPW <= 8, "tiles per wave");
  constexpr int NP = kWaves * 32 * TPW;
  constexpr int NB = NP / 16;
  constexpr int NPW = NP / 32;
  __shared__ uint32_t dring[kRing][kRT];
  __shared__ uint4 lutp[256];
  __shared__ uint4 lutn[256];
  __shared__ __align__(16) float fblk[32][kFT];  // the owner's tile, column-major
  __shared__ float thr[kTRing][kRT][17];
  __shared__ uint32_t xs[kRT][NPW + 1];
  __shared__ double es[kRT];
  __shared__ double ebs[kRT];
  __shared__ __align__(16) uint4 qds[64];  // the owner's fp32 16x16 diagonal sub-block
  __shared__ int d_count;
  __shared__ int thr_cnt[kTRing];
  __shared__ int abort_flag;
  __shared__ float tsl[kRT];  // per-replica temperature scales (tsr), if given

  const int tid = threadIdx.x;
  const int w = tid >> 6;
  const int lane = tid & 63;
  const int lc = lane & 31;
  const int kh = lane >> 5;
  // gpi work-groups per instance: this one anneals replicas [rloc, rloc+64)
  // of instance inst, stored at rows [rbase, rbase+64) of the state arrays
  const int inst = blockIdx.x / gpi;
  const uint32_t rbase = blockIdx.x * kRT;
  const uint32_t rloc = (uint32_t)(blockIdx.x % gpi) * kRT;
  const uint32_t tag = ((inst_base + (uint32_t)inst) << 8) | TAG_SWEEP_THR;
  qpack += (size_t)inst * NB * NP * 4;
  qdb += (size_t)inst * NB * 256;
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
    tsl[tid] = tsr ? tsr[rbase + tid] : 0.0f;
  }
  if (tid < kTRing) thr_cnt[tid] = kThrWriters;
  if (tid == 0) {
    d_count = 0;
    abort_flag = 0;
  }
  // accumulator element i of lane (lc, kh): replica 32m + 8(i/4) + 4kh + i%4,
  // column tile_col(w, tt, lc)
  f32x16 C[2][TPW];
#pragma unroll
  for (int m = 0; m < 2; ++m)
#pragma unroll
    for (int tt = 0; tt < TPW; ++tt)
#pragma unroll
      for (int i = 0; i < 16; ++i)
        C[m][tt][i] = fields[(size_t)(rbase + 32 * m + 8 * (i >> 2) + 4 * kh + (i & 3)) * NP +
                             tile_col(w, tt, lc)];
  // thresholds of the first kTRing blocks (256 items each); block g >= kTRing
  // is computed by the three non-owners of step g - kLA
  for (int it = tid; it < kTRing * 256; it += 256) {
    const int gblk = it >> 8;
    if (gblk < G)
      block_thresholds(it & 255, gblk % NB, gblk / NB, n, rloc, sweep0, k0, k1, tag,
                       tsr ? tsr[rbase + ((it & 255) >> 2)] : tscale[gblk / NB], thr[gblk]);
  }
  __syncthreads();

  // B operands (rows of block x, all tiles): the next D block to apply
  uint4 Bn[TPW][2];
  auto load_ops = [&](int x) {
#pragma unroll
    for (int tt = 0; tt < TPW; ++tt)
#pragma unroll
      for (int p = 0; p < 2; ++p)
        Bn[tt][p] = qpack[((size_t)x * NP + tile_col(w, tt, lc)) * 4 + 2 * p + kh];
  };
  // fetched a step early for this wave's next owner step: its tile's operands
  // and its diagonal sub-block
  uint4 Bo[2] = {make_uint4(0u, 0u, 0u, 0u), make_uint4(0u, 0u, 0u, 0u)};
  uint4 qdn = make_uint4(0u, 0u, 0u, 0u);
  if (w == 0) qdn = reinterpret_cast<const uint4 *>(qdb)[lane];  // block 0 (step 0)
  load_ops(0);

  bool ok = true;
  for (int g = 0; g < G && ok; ++g) {
    const int b = g % NB;
    const int owner = b & 3;
    const int b1 = (g + 1) % NB;
    bf16x8 A[2];
    if (g > 0) {
      ok = wait_ge(&d_count, g, &abort_flag);  // D_{g-1} published
      if (!ok) break;
      load_a(dring[(g - 1) % kRing], lutp, lutn, lc, kh, A);
    }
    // the owner of the next step is one of this step's helpers; its loads sit
    // in the helper branch, so the owner branch waits for none of them
    const bool next_owner = (b1 & 3) == w && g + 1 < G;
    if (owner == w) {
      const int k = b >> 2;
      const int t = k >> 1;     // this wave's tile holding block b
      const int half = k & 1;   // its lanes 16*half .. +15
      qds[lane] = qdn;
      const double e_prev = es[lane];
      // (a) bring tile t up to date and transpose it (static indices, uniform branch)
#pragma unroll
      for (int tt = 0; tt < TPW; ++tt) {
        if (tt != t) continue;
        if (g > 0) {
#pragma unroll
          for (int p = 0; p < 2; ++p) {
            const bf16x8 Bt = as_bf16x8(Bo[p]);
#pragma unroll
            for (int m = 0; m < 2; ++m) C[m][tt] = mfma32(A[m], Bt, C[m][tt]);
          }
        }
#pragma unroll
        for (int m = 0; m < 2; ++m)
#pragma unroll
          for (int q = 0; q < 4; ++q)
            *reinterpret_cast<float4 *>(&fblk[lc][32 * m + 8 * q + 4 * kh]) =
                make_float4(C[m][tt][4 * q], C[m][tt][4 * q + 1], C[m][tt][4 * q + 2],
                            C[m][tt][4 * q + 3]);
      }
      wave_sync();
      // ---- sequential decisions of block b, lane = replica ----
      const float(*tb_thr)[17] = thr[g % kTRing];
      const float *fcol = &fblk[16 * half][lane];
      const uint32_t word = xs[lane][b >> 1];
      const int sh = (b & 1) * 16;
      uint32_t xb = (word >> sh) & 0xFFFFu;
      const uint32_t x0 = xb;
      float desum = 0.0f;
      // The block's 16 fields and thresholds of this lane's replica are read
      // at once; c[j] starts as the field of spin j and gains the in-block
      // corrections.  Sub-block rows (row == column: the diagonal block of a
      // symmetric Q is symmetric) are read two steps ahead; sched_barrier
      // keeps the compiler from hoisting all 16 steps' reads.
      float c[16], th[16];
#pragma unroll
      for (int jj = 0; jj < 16; ++jj) {
        c[jj] = fcol[jj * kFT];
        th[jj] = tb_thr[lane][jj];
      }
      // qds through a base register, so every row read is one ds_read with an
      // immediate offset (a constant address costs a v_mov per read)
      uint32_t qa = lds_addr(qds);
      asm volatile("" : "+v"(qa));
      const __attribute__((address_space(3))) f32x4 *qdv =
          (const __attribute__((address_space(3))) f32x4 *)(uintptr_t)qa;
      f32x4 qn[2][4];
#pragma unroll
      for (int r = 0; r < 2; ++r)
#pragma unroll
        for (int q = 0; q < 4; ++q)
          if (4 * q + 3 > r) qn[r][q] = qdv[r * 4 + q];
#pragma unroll
      for (int jj = 0; jj < 16; ++jj) {
        float qc[16];
#pragma unroll
        for (int q = 0; q < 4; ++q)
          if (4 * q + 3 > jj) {
            qc[4 * q] = qn[jj & 1][q][0];
            qc[4 * q + 1] = qn[jj & 1][q][1];
            qc[4 * q + 2] = qn[jj & 1][q][2];
            qc[4 * q + 3] = qn[jj & 1][q][3];
          }
        if (jj < 14) {
#pragma unroll
          for (int q = 0; q < 4; ++q)
            if (4 * q + 3 > jj + 2) qn[jj & 1][q] = qdv[(jj + 2) * 4 + q];
        }
        // de = (1 - 2 x_j) h and d = +-1 (or 0) by the sign bit of x_j; d * h
        // is exact, so desum gains exactly de on a flip
        const float h = c[jj];
        const uint32_t sg = (x0 << (31 - jj)) & 0x80000000u;
        const float de = __uint_as_float(__float_as_uint(h) ^ sg);
        const bool acc = de < th[jj];
        const float d = acc ? __uint_as_float(0x3F800000u | sg) : 0.0f;
        xb ^= acc ? (1u << jj) : 0u;
        desum = fmaf(d, h, desum);
#pragma unroll
        for (int mm = jj + 1; mm < 16; ++mm) c[mm] = fmaf(qc[mm], d, c[mm]);
        __builtin_amdgcn_sched_barrier(0);
      }
      xs[lane][b >> 1] = (word & ~(0xFFFFu << sh)) | (xb << sh);
      const uint32_t flips = xb ^ x0;
      const uint32_t dmask = (flips & ~x0) | ((flips & x0) << 16);
      dring[g % kRing][lane] = dmask;
      const double e_now = e_prev + (double)desum;
      es[lane] = e_now;
      // end of a sweep: best-state tracking (before the next owner may change xs)
      if (b == NB - 1 && e_now < ebs[lane]) {
        ebs[lane] = e_now;
        uint32_t *dst = bestbits + (size_t)(rbase + lane) * NPW;
        for (int q = 0; q < NPW; ++q) dst[q] = xs[lane][q];
      }
      if (lane == 0) lds_store(&d_count, g + 1);
      // (b) after publishing: D_{g-1} into the other tiles
      if (g > 0) {
#pragma unroll
        for (int p = 0; p < 2; ++p)
#pragma unroll
          for (int tt = 0; tt < TPW; ++tt) {
            if (tt == t) continue;
            const bf16x8 Bt = as_bf16x8(Bn[tt][p]);
#pragma unroll
            for (int m = 0; m < 2; ++m) C[m][tt] = mfma32(A[m], Bt, C[m][tt]);
          }
      }
    } else {
      if (next_owner) {
        // next step is this wave's: its tile of block b1 needs D_g (rows of b)
        const int t1 = (b1 >> 2) >> 1;
#pragma unroll
        for (int p = 0; p < 2; ++p)
          Bo[p] = qpack[((size_t)b * NP + tile_col(w, t1, lc)) * 4 + 2 * p + kh];
        qdn = reinterpret_cast<const uint4 *>(qdb)[(size_t)b1 * 64 + lane];
      }
      if (g > 0) {
        const int hw = (w - owner - 1 + kWaves) % kWaves;  // 0 .. kHelpers-1
        if (hw < kHelpers - kThrWriters) {
          // an owner of one of the next steps: D_{g-1} into every tile only
#pragma unroll
          for (int p = 0; p < 2; ++p)
#pragma unroll
            for (int tt = 0; tt < TPW; ++tt) {
              const bf16x8 Bt = as_bf16x8(Bn[tt][p]);
#pragma unroll
              for (int m = 0; m < 2; ++m) C[m][tt] = mfma32(A[m], Bt, C[m][tt]);
            }
        } else {
          // D_{g-1} into every tile, with the thresholds of block g+kLA
          // computed under the MFMAs (their slot was last read in step
          // g+kLA-kTRing < g): items item + 64i, i < kThrItems (replica
          // r = item/4, spins 16*bn + 4q .. +3, q = item%4); their 10 Philox
          // rounds and 4 logarithms are spread over the MFMA pairs
          const int g2 = min(g + kLA, G - 1);
          const int item = (hw - (kHelpers - kThrWriters)) * 64 * kThrItems + lane;
          float tsi[kThrItems];
          {
            const float ts2 = tsr ? 0.0f : tscale[g2 / NB];
#pragma unroll
            for (int i = 0; i < kThrItems; ++i) tsi[i] = tsr ? tsl[(item + 64 * i) >> 2] : ts2;
          }
          const int j0 = (g2 % NB) * 16 + 4 * (item & 3);
          uint4 ph[kThrItems];
#pragma unroll
          for (int i = 0; i < kThrItems; ++i)
            ph[i] = make_uint4((uint32_t)(j0 >> 2), rloc + (uint32_t)((item + 64 * i) >> 2),
                               (uint32_t)(sweep0 + g2 / NB), tag);
          uint32_t pk0 = k0, pk1 = k1;
          float tv[kThrItems][4];
#pragma unroll
          for (int p = 0; p < 2; ++p)
#pragma unroll
            for (int tt = 0; tt < TPW; ++tt) {
              const bf16x8 Bt = as_bf16x8(Bn[tt][p]);
#pragma unroll
              for (int m = 0; m < 2; ++m) C[m][tt] = mfma32(A[m], Bt, C[m][tt]);
              const int slot = p * TPW + tt;
#pragma unroll
              for (int u = 0; u < 14; ++u) {
                if ((u * 2 * TPW) / 14 != slot) continue;
                if (u < 10) {
                  if (u) {
                    pk0 += kPhiloxW0;
                    pk1 += kPhiloxW1;
                  }
#pragma unroll
                  for (int i = 0; i < kThrItems; ++i) philox_round(ph[i], pk0, pk1, 0);
                } else {
#pragma unroll
                  for (int i = 0; i < kThrItems; ++i) {
                    const float x = neg_ln_u(pick(ph[i], u - 10)) * tsi[i];
                    tv[i][u - 10] = j0 + (u - 10) < n ? x : -INFINITY;
                  }
                }
              }
            }
          if (g + kLA < G) {
            const int sl = (g + kLA) % kTRing;
#pragma unroll
            for (int i = 0; i < kThrItems; ++i) store_thresholds(item + 64 * i, tv[i], thr[sl]);
            if (lane == 0) lds_add(&thr_cnt[sl]);
          }
        }
      }
      // Bo and qdn have arrived long since.  The role branches are divergent
      // to the compiler, which lays the owner branch out after this one; this
      // wait, on every path through this branch, keeps it from waiting for
      // them (and for the next operands behind them) at the owner's MFMAs.
      __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0)
    }
    if (next_owner) {
      // off the critical path: the thresholds of the next owner step
      ok = wait_ge(&thr_cnt[(g + 1) % kTRing], kThrWriters * ((g + 1) / kTRing + 1), &abort_flag);
      if (!ok) break;
    }
    // D_0 .. D_{g-1} are in every tile of this wave; fetch D_g's operands
    load_ops(b);
  }
  // drain: D_{G-1} into every tile
  if (ok && G > 0) {
    ok = wait_ge(&d_count, G, &abort_flag);
    if (ok) {
      bf16x8 A[2];
      load_a(dring[(G - 1) % kRing], lutp, lutn, lc, kh, A);
#pragma unroll
      for (int p = 0; p < 2; ++p)
#pragma unroll
        for (int tt = 0; tt < TPW; ++tt) {
          const bf16x8 Bt = as_bf16x8(Bn[tt][p]);
#pragma unroll
          for (int m = 0; m < 2; ++m) C[m][tt] = mfma32(A[m], Bt, C[m][tt]);
        }
    }
  }
  if (!ok && lane == 0) lds_store(&abort_flag, 1);
#pragma unroll
  for (int m = 0; m < 2; ++m)
#pragma unroll
    for (int tt = 0; tt < TPW; ++tt)
#pragma unroll
      for (int i = 0; i < 16; ++i)
        fields[(size_t)(rbase + 32 * m + 8 * (i >> 2) + 4 * kh + (i & 3)) * NP +
               tile_col(w, tt, lc)] = C[m][tt][i];
  __syncthreads();
  for (int i = tid; i < kRT * NPW; i += 256)
    xbits[(size_t)(rbase + i / NPW) * NPW + i % NPW] = xs[i / NPW][i % NPW];
  if (tid < kRT) {
    energy[rbase + tid] = es[tid];
    ebest[rbase + tid] = ebs[tid];
  }
  if (tid == 0 && abort_flag) atomicOr(status, 1);
}

// Random initial bits of instance i (rows [i*R, (i+1)*R)): spin j of replica
// r from counter (j>>7, r, 0, ((instance + i) << 8) | TAG_SWEEP_INIT).
__global__ __launch_bounds__(256) void sweep_init_kernel(int n, int npw, int R, int I, uint32_t k0,
                                                         uint32_t k1, uint32_t instance,
                                                         uint32_t *__restrict__ xbits) {
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (size_t)I * R * npw) return;
  const size_t row = idx / npw;
  const uint32_t inst = (uint32_t)(row / R);
  const uint32_t r = (uint32_t)(row % R);
  const int wd = (int)(idx % npw);
  const int j0 = wd * 32;
  const uint32_t tag = ((instance + inst) << 8) | TAG_SWEEP_INIT;
  const uint4 o = philox4x32_10(make_uint4((uint32_t)(j0 >> 7), r, 0u, tag), k0, k1);
  uint32_t v = pick(o, (j0 >> 5) & 3);
  const int valid = n - j0;
  if (valid < 32) v = valid <= 0 ? 0u : (v & ((1u << valid) - 1u));
  xbits[idx] = v;
}

// Warm start of instance i: x_j = (u < p[i][j]), u = (w >> 8) * 2^-24 from
// counter (j>>2, r, 0, ((instance + i) << 8) | TAG_WARM).
__global__ __launch_bounds__(256) void warm_init_kernel(int n, int npw, int R, int I, uint32_t k0,
                                                        uint32_t k1, uint32_t instance,
                                                        const float *__restrict__ p,
                                                        uint32_t *__restrict__ xbits) {
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (size_t)I * R * npw) return;
  const size_t row = idx / npw;
  const uint32_t inst = (uint32_t)(row / R);
  const uint32_t r = (uint32_t)(row % R);
  const int wd = (int)(idx % npw);
  const float *pi = p + (size_t)inst * n;
  const uint32_t tag = ((instance + inst) << 8) | TAG_WARM;
  uint32_t v = 0;
  for (int q = 0; q < 8; ++q) {
    const int j0 = wd * 32 + 4 * q;
    if (j0 >= n) break;
    const uint4 o = philox4x32_10(make_uint4((uint32_t)(j0 >> 2), r, 0u, tag), k0, k1);
    for (int jj = 0; jj < 4; ++jj) {
      const int j = j0 + jj;
      if (j < n && (float)(pick(o, jj) >> 8) * 0x1p-24f < pi[j]) v |= 1u << (4 * q + jj);
    }
  }
  xbits[idx] = v;
}

// Population annealing: row (instance i, r) of the new population is row
// (i, anc[i][r]) of the old one for r < R; thread = (row, column).
__global__ __launch_bounds__(256) void dense_gather_kernel(int Rp, int R, int npw, int np,
                                                           const uint32_t *__restrict__ anc,
                                                           const uint32_t *__restrict__ x,
                                                           const float *__restrict__ f,
                                                           const double *__restrict__ e,
                                                           uint32_t *__restrict__ x2,
                                                           float *__restrict__ f2,
                                                           double *__restrict__ e2) {
  const int row = blockIdx.y;
  const int col = blockIdx.x * 256 + threadIdx.x;
  const int inst = row / Rp, r = row % Rp;
  const int src = r < R ? inst * Rp + (int)anc[(size_t)inst * R + r] : row;
  if (col < np) f2[(size_t)row * np + col] = f[(size_t)src * np + col];
  if (col < npw) x2[(size_t)row * npw + col] = x[(size_t)src * npw + col];
  if (col == 0) e2[row] = e[src];
}

void launch_dense_gather(hipStream_t st, int num_instances, int Rp, int R, int npw, int np,
                         const uint32_t *anc, const uint32_t *x, const float *f, const double *e,
                         uint32_t *x2, float *f2, double *e2) {
  const dim3 grid((unsigned)((np + 255) / 256), (unsigned)(num_instances * Rp));
  dense_gather_kernel<<<grid, 256, 0, st>>>(Rp, R, npw, np, anc, x, f, e, x2, f2, e2);
  OS_HIP_CHECK(hipGetLastError());
}

// Replica exchange round `round` on the GPU, for the dense and the sparse
// paths (replica r of instance i at row i * Rp + r): thread = (instance,
// ladder, pair p); slots k = round % 2 + 2p and k + 1 of the ladder swap
// their replicas a, b iff d = (1/ts_k - 1/ts_{k+1}) (E_a - E_b) >= 0 or
// u < exp(d), u = (w >> 8) 2^-24 from word 0 of Philox counter (k, l, round,
// ((instance_base + i) << 8) | TAG_EXCHANGE); the swapped replicas take their
// new slots' scales in tsr.  The pairs of a round are disjoint.
__global__ __launch_bounds__(64) void exchange_round_kernel(
    int I, int L, int K, int R, int Rp, uint32_t round, uint32_t k0, uint32_t k1,
    uint32_t inst_base, const float *__restrict__ ts_slots, const double *__restrict__ e,
    uint32_t *__restrict__ rep_at, uint32_t *__restrict__ slot_of, float *__restrict__ tsr) {
  const int pairs = K / 2;
  const int idx = blockIdx.x * 64 + threadIdx.x;
  if (idx >= I * L * pairs) return;
  const int i = idx / (L * pairs), l = (idx / pairs) % L, p = idx % pairs;
  const int k = (int)(round & 1u) + 2 * p;
  if (k + 1 >= K) return;
  const size_t ro = (size_t)i * Rp;
  uint32_t *ra = rep_at + (size_t)i * R + (size_t)l * K;
  const uint32_t a = ra[k], b = ra[k + 1];
  const double d =
      (1.0 / (double)ts_slots[k] - 1.0 / (double)ts_slots[k + 1]) * (e[ro + a] - e[ro + b]);
  const uint4 o = philox4x32_10(
      make_uint4((uint32_t)k, (uint32_t)l, round, ((inst_base + (uint32_t)i) << 8) | TAG_EXCHANGE),
      k0, k1);
  const double u = (double)(o.x >> 8) * 0x1p-24;
  if (d >= 0.0 || u < exp(d)) {
    ra[k] = b;
    ra[k + 1] = a;
    slot_of[(size_t)i * R + a] = (uint32_t)(k + 1);
    slot_of[(size_t)i * R + b] = (uint32_t)k;
    tsr[ro + a] = ts_slots[k + 1];
    tsr[ro + b] = ts_slots[k];
  }
}

void launch_exchange_round(hipStream_t st, int I, int L, int K, int R, int Rp, uint32_t round,
                           uint64_t seed, uint32_t inst_base, const float *ts_slots,
                           const double *e, uint32_t *rep_at, uint32_t *slot_of, float *tsr) {
  const int threads = I * L * (K / 2);
  if (threads == 0) return;
  exchange_round_kernel<<<(threads + 63) / 64, 64, 0, st>>>(
      I, L, K, R, Rp, round, (uint32_t)seed, (uint32_t)(seed >> 32), inst_base, ts_slots, e,
      rep_at, slot_of, tsr);
  OS_HIP_CHECK(hipGetLastError());
}

int dense_max_n() { return kWaves * 32 * 8; }

// Column tiles (32 columns) per wave are instantiated for these counts; n is
// padded up to the next one (128 columns per step).
static const int kTpwList[] = {1, 2, 3, 4, 6, 8};

int dense_padded_n(int n) {
  const int need = (n + kWaves * 32 - 1) / (kWaves * 32);
  for (int t : kTpwList)
    if (t >= need) return t * kWaves * 32;
  return -1;
}

void launch_dense_sweeps(hipStream_t st, int n, int np, int num_instances, const uint4 *qpack,
                         const float *qdb, const float *tscale, const float *tsr, int sweep0,
                         int nsweeps, uint64_t seed, uint32_t instance, int R_pad, float *fields,
                         uint32_t *xbits, uint32_t *bestbits, double *energy, double *ebest,
                         int *status) {
  const uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
  const int gpi = R_pad / kRT;
  const dim3 grid(num_instances * gpi), block(kWaves * 64);
  switch (np / (kWaves * 32)) {
#define OS_DENSE_CASE(T)                                                                    \
  case T:                                                                                   \
    dense_sweep_kernel<T><<<grid, block, 0, st>>>(n, qpack, qdb, tscale, tsr, sweep0, nsweeps, k0, \
                                                  k1, instance, gpi, fields, xbits, bestbits, \
                                                  energy, ebest, status);                   \
    break;
    OS_DENSE_CASE(1)
    OS_DENSE_CASE(2)
    OS_DENSE_CASE(3)
    OS_DENSE_CASE(4)
    OS_DENSE_CASE(6)
    OS_DENSE_CASE(8)
#undef OS_DENSE_CASE
    default:
      throw Error(-3, "dense sweep kernel: unsupported padded size");
  }
  OS_HIP_CHECK(hipGetLastError());
}

void launch_sweep_init(hipStream_t st, int n, int npw, int R, uint64_t seed, uint32_t instance,
                       uint32_t *xbits, int num_instances) {
  const size_t total = (size_t)num_instances * R * npw;
  if (total == 0) return;
  sweep_init_kernel<<<(unsigned)((total + 255) / 256), 256, 0, st>>>(
      n, npw, R, num_instances, (uint32_t)seed, (uint32_t)(seed >> 32), instance, xbits);
  OS_HIP_CHECK(hipGetLastError());
}

void launch_warm_init(hipStream_t st, int n, int npw, int R, uint64_t seed, uint32_t instance,
                      const float *p, uint32_t *xbits, int num_instances) {
  const size_t total = (size_t)num_instances * R * npw;
  if (total == 0) return;
  warm_init_kernel<<<(unsigned)((total + 255) / 256), 256, 0, st>>>(
      n, npw, R, num_instances, (uint32_t)seed, (uint32_t)(seed >> 32), instance, p, xbits);
  OS_HIP_CHECK(hipGetLastError());
}

}  // namespace os
